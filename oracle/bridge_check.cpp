// bridge_check.cpp -- TEST INFRASTRUCTURE: drives the reference's own engine (compiled from its sources in place)
// through the drop-ins of insiliscope_amd/integration (SplatGpuIscg.cpp in place of Sim/SplatGpu.cpp, GpuSimIscg.cpp
// in place of Sim/GpuSimD3D11.cpp, WidefieldIscgHost.h over the unmodified WidefieldAccelerator), and returns the
// GPU-path results beside the same engine's CPU results. Built by oracle/Makefile (target `bridge`) into
// _ref/libiscbridge.so, linked against libiscg.so; only tests/ load it.
#include "GpuSimD3D11.h"
#include "PsfGeneratorBridge.h"
#include "SMLMNoise.h"
#include "SMLMSimulation.h"
#include "ScopeMovie.h"
#include "SplatGpu.h"
#include "WidefieldRender.h"
#include "Illumination.h"

#include "../insiliscope_amd/integration/WidefieldIscgHost.h"

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_err;

sim::PsfKernelCache Kernel(const float* planes, int nz, int size, int os, double zStepNm, int interp)
{
   sim::PsfKernelCache c;
   c.oversampling = os;
   c.halfWidthOversampled = (size - 1) / 2;
   c.sizeOversampled = size;
   c.nz = nz;
   c.zStepNm = zStepNm;
   c.interpMode = static_cast<sim::PsfInterpMode>(interp);
   c.blockSumWidth = size + os - 1;
   sim::PsfKernelPlanes d;
   const size_t plane = static_cast<size_t>(size) * size;
   d.planes.resize(static_cast<size_t>(nz));
   for (int z = 0; z < nz; ++z)
      d.planes[static_cast<size_t>(z)].assign(planes + z * plane, planes + (z + 1) * plane);
   c.SetData(std::move(d));
   c.valid = true;
   return c;
}

std::vector<sim::BlinkEvent> Events(const double* ev6, int n)
{
   std::vector<sim::BlinkEvent> out(static_cast<size_t>(n));
   for (int i = 0; i < n; ++i)
   {
      const double* e = ev6 + 6 * i;
      sim::BlinkEvent& b = out[static_cast<size_t>(i)];
      b.xUm = e[0];
      b.yUm = e[1];
      b.zNm = e[2];
      b.tStart = e[3];
      b.tEnd = e[4];
      b.brightness = e[5];
   }
   return out;
}

sim::CameraNoiseParams Camera(const double* c)
{
   // quantumEfficiency, darkCurrentElectrons, gainPhotonsPerAdu, readNoiseElectrons, emccd, emGain, cicElectrons,
   // bitDepth
   sim::CameraNoiseParams p;
   p.quantumEfficiency = c[0];
   p.darkCurrentElectrons = c[1];
   p.gainPhotonsPerAdu = c[2];
   p.readNoiseElectrons = c[3];
   p.emccd = c[4] != 0.0;
   p.emGain = c[5];
   p.cicElectrons = c[6];
   p.bitDepth = static_cast<int>(c[7]);
   return p;
}

}  // namespace

extern "C" {

const char* bridge_error() { return g_err.c_str(); }

// The reference's GPU record path frame by frame: CollectGpuEmitters (Sim/SMLMSimulation.cpp:219) of the events,
// then via = 0: sim::SplatGpuHost::RenderFrames (this build's SplatGpuIscg.cpp, made by MakeSplatGpuHost with its
// self-check), via = 1: sim::GpuSimulator::RenderFrames (GpuSimIscg.cpp) -> gpu; and the reference's CPU frames,
// RenderPhotonImage + ApplyNoiseChain with the same offset map -> cpu (nFrames x h x w ADU each). 0 or -1.
int32_t bridge_splat_frames(int32_t via, const double* ev6, int32_t nEv, int32_t nFrames, uint32_t w, uint32_t h,
                            double pixelNm, double perBlink, double bg, const float* planes, int32_t nz, int32_t size,
                            int32_t os, double zStepNm, int32_t interp, const double* cam6, double offsetAdu,
                            uint32_t seed, uint16_t* gpu, uint16_t* cpu)
{
   const sim::PsfKernelCache k = Kernel(planes, nz, size, os, zStepNm, interp);
   const std::vector<sim::BlinkEvent> evs = Events(ev6, nEv);
   const sim::CameraNoiseParams cam = Camera(cam6);
   const size_t n = static_cast<size_t>(w) * h;
   const std::vector<float> offset(n, static_cast<float>(offsetAdu)), none;
   std::vector<std::vector<sim::GpuSplatEmitter>> ems(static_cast<size_t>(nFrames));
   std::vector<uint32_t> frames;
   for (long f = 0; f < nFrames; ++f)
   {
      sim::CollectGpuEmitters(evs, f, w, h, pixelNm, perBlink, 0.0, 0.0, k, 0.0, nullptr, ems[static_cast<size_t>(f)]);
      frames.push_back(static_cast<uint32_t>(f));
   }
   std::vector<std::vector<uint16_t>> got(static_cast<size_t>(nFrames));
   std::vector<std::vector<uint16_t>*> outs;
   for (auto& g : got)
      outs.push_back(&g);
   std::string err;
   if (via == 0)
   {
      std::string info;
      std::unique_ptr<sim::SplatGpuHost> host = sim::MakeSplatGpuHost(nullptr, info);
      if (!host)
      {
         g_err = info;
         return -1;
      }
      std::vector<std::vector<float>> extra(static_cast<size_t>(nFrames), std::vector<float>(n, static_cast<float>(bg)));
      std::vector<const std::vector<float>*> ex;
      for (const auto& e : extra)
         ex.push_back(&e);
      if (!host->SetKernels({&k}, err) || !host->SetStatic(w, h, 0, offset, none, none, cam, 0, err) ||
          !host->RenderFrames(ems, frames, ex, cam, seed, outs, err))
      {
         g_err = err;
         return -1;
      }
   }
   else
   {
      std::string info;
      std::unique_ptr<sim::GpuSimulator> sim = sim::GpuSimulator::Create(info);
      if (!sim)
      {
         g_err = info;
         return -1;
      }
      const std::vector<double> scales(static_cast<size_t>(nFrames), 1.0);
      if (!sim->SetKernels({&k}, err) || !sim->SetStatic(w, h, offset, none, none, none, bg, cam, err) ||
          !sim->RenderFrames(ems, frames, scales, cam, seed, outs, err))
      {
         g_err = err;
         return -1;
      }
   }
   sim::PixelOffsetMap om;
   om.width = w;
   om.height = h;
   om.offset = offset;
   for (long f = 0; f < nFrames; ++f)
   {
      std::vector<float> img;
      sim::RenderPhotonImage(img, w, h, evs, f, pixelNm, 1.0, perBlink, bg, 0.0, 0.0, &k);
      std::vector<uint16_t> want;
      sim::ApplyNoiseChain(img, want, w, h, cam, om, sim::PixelGainMap(), sim::PixelReadNoiseMap(), seed,
                           static_cast<uint32_t>(f));
      const auto& g = got[static_cast<size_t>(f)];
      if (g.size() != n || want.size() != n)
      {
         g_err = "frame size";
         return -1;
      }
      std::memcpy(gpu + f * n, g.data(), n * sizeof(uint16_t));
      std::memcpy(cpu + f * n, want.data(), n * sizeof(uint16_t));
   }
   return 0;
}

// A whole movie through the engine (RenderScopeMovie, Sim/ScopeMovie.h:477, as insiliscope_cli runs it): the spec's
// options as names and values, gpu = 1 with this build's SplatGpuHost installed as the process-wide blink splat
// (SetDefaultBlinkSplatGpu), 0 with none (the CPU). out: frames x height x width ADU (capacity outCap pixels);
// info[0..3] = width, height, frames, blinks; text: the GPU host's info line. Returns the frames written or -1.
int64_t bridge_movie(const char* const* names, const double* values, int32_t nOpts, int32_t gpu, uint16_t* out,
                     uint64_t outCap, int64_t* info, char* text, int32_t textCap)
{
   sim::ScopeSpec spec;
   for (int i = 0; i < nOpts; ++i)
      if (!sim::ScopeSpecSet(spec, names[i], values[i]))
      {
         g_err = std::string("unknown option ") + names[i];
         return -1;
      }
   if (gpu)
      sim::SetDefaultBlinkSplatGpu([](std::string& inf) { return sim::MakeSplatGpuHost(nullptr, inf); });
   else
      sim::SetDefaultBlinkSplatGpu({});
   sim::ScopeMovieInfo mi;
   std::string err;
   int64_t written = 0;
   const bool ok = sim::RenderScopeMovie(
      spec,
      [&](long f, const std::vector<uint16_t>& adu) {
         if (static_cast<uint64_t>(f + 1) * adu.size() > outCap)
            return false;
         std::memcpy(out + static_cast<size_t>(f) * adu.size(), adu.data(), adu.size() * sizeof(uint16_t));
         written = f + 1;
         return true;
      },
      mi, err);
   const std::string hostInfo = gpu ? sim::DefaultBlinkSplatGpuInfo() : std::string("cpu");
   if (gpu && !sim::DefaultBlinkSplatGpu())
   {
      g_err = "the GPU splat host was not made: " + hostInfo;
      return -1;
   }
   sim::SetDefaultBlinkSplatGpu({});
   if (!ok)
   {
      g_err = err;
      return -1;
   }
   info[0] = mi.width;
   info[1] = mi.height;
   info[2] = mi.frames;
   info[3] = static_cast<int64_t>(mi.blinks);
   if (text && textCap > 0)
   {
      std::strncpy(text, hostInfo.c_str(), static_cast<size_t>(textCap) - 1);
      text[textCap - 1] = 0;
   }
   return written;
}

// A GPU-mode WidefieldScene (the reference's own, Sim/WidefieldRender.h) of a persistent population (density
// [nzD][ny][nx] at dzNm per plane, flat illumination, the FOV moved (dxUm, dyUm) off the cell edges) rendered at the
// foci through WidefieldIscgHost -> gpu, and the same scene without an accelerator -> cpu (nF x ny x nx each).
// *used = 1 when the scene served its foci from the accelerator. 0 or -1.
int32_t bridge_wf_scene(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzNm, const float* planes,
                        int32_t nzP, int32_t size, int32_t os, double zStepNm, double kernelCapUm,
                        const double* focusUm, int32_t nF, double dxUm, double dyUm, float* gpu, float* cpu,
                        int32_t* used)
{
   const sim::PsfKernelCache c = Kernel(planes, nzP, size, os, zStepNm, 0);
   sim::KernelWidefieldPsf psf(c, 1);
   sim::WidefieldSceneSpec spec;
   spec.pixelUm = 0.1;
   spec.originXUm = 0.05 + dxUm;
   spec.originYUm = 0.05 + dyUm;
   spec.width = static_cast<unsigned>(nx);
   spec.height = static_cast<unsigned>(ny);
   spec.focusWorldUm = focusUm[0];
   spec.slabHalfUm = 0.0;
   spec.grid.upscale = 1;
   spec.grid.zPlaneNm = dzNm;
   spec.marginUm = 0.0;
   spec.kernelCapUm = kernelCapUm;
   spec.eta = 1.0;
   spec.unitDose = 1.0;
   sim::FlatIllumination ill(nx * 0.1, ny * 0.1);
   unsigned fx = 0, fy = 0;
   double frx = 0, fry = 0;
   const sim::WidefieldGridSpec g = sim::WidefieldScene::GridSpecFor(ill, spec, &fx, &fy, &frx, &fry);
   sim::WidefieldDyeGrid G;
   G.spec = g;
   G.k0 = 0;
   G.nz = static_cast<unsigned>(nzD);
   const size_t gp = static_cast<size_t>(g.nx) * g.ny;
   G.persistent.assign(gp * nzD, 0.0f);
   double total = 0.0;
   for (int z = 0; z < nzD; ++z)
      for (int y = 0; y < ny; ++y)
         for (int x = 0; x < nx; ++x)
         {
            const float v = density[(static_cast<size_t>(z) * ny + y) * nx + x];
            G.persistent[z * gp + static_cast<size_t>(fy + y) * g.nx + fx + x] = v;
            total += v;
         }
   G.nPersistent = static_cast<long>(total);
   iscg_context* h = nullptr;
   if (iscg_create(0, &h) != 0)
   {
      g_err = iscg_last_error(nullptr);
      return -1;
   }
   int32_t rc = 0;
   {
      sim::WidefieldIscgHost host(h);
      sim::WidefieldScene gs, cs;
      gs.SetGpuMode(true);
      cs.SetGpuMode(true);
      gs.SetAccelerator(&host);
      std::string err;
      const std::vector<double> foci(focusUm, focusUm + nF);
      std::vector<sim::WidefieldImages> gi, ci;
      if (!gs.UpdateFromGrid(G, ill, spec, psf, err) || !cs.UpdateFromGrid(G, ill, spec, psf, err) ||
          !gs.FocusSeries(foci, gi, err) || !cs.FocusSeries(foci, ci, err))
      {
         g_err = err;
         rc = -1;
      }
      else
      {
         *used = gs.UsingAccelerator() ? 1 : 0;
         const size_t n = static_cast<size_t>(nx) * ny;
         for (int f = 0; f < nF && rc == 0; ++f)
         {
            const auto& a = gi[static_cast<size_t>(f)].persistent;
            const auto& b = ci[static_cast<size_t>(f)].persistent;
            if (a.size() != n || b.size() != n)
            {
               g_err = "image size";
               rc = -1;
               break;
            }
            std::memcpy(gpu + f * n, a.data(), n * sizeof(float));
            std::memcpy(cpu + f * n, b.data(), n * sizeof(float));
         }
      }
   }
   iscg_destroy(h);
   return rc;
}

}  // extern "C"
