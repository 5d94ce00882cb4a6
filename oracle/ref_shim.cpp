// ref_shim.cpp -- extern "C" entry points over the reference's own CPU functions, compiled together with the
// reference sources (read in place under the reference checkout) into oracle/_ref/libiscref.so.
//
// TEST INFRASTRUCTURE ONLY: used to write golden fixtures (tests/golden/make_golden.py), to cross-check the FP64
// restatement in oracle.c, and as the timed CPU baseline. Nothing here is linked into the product library.
//
// Each entry point forwards to the named reference function with the arguments spelled out; the only code of its
// own is argument marshalling.

#include "PsfGeneratorBridge.h"
#include "SMLMNoise.h"
#include "SMLMSimulation.h"
#include "SMLMZernike.h"
#include "WidefieldRender.h"
#include "Drift.h"
#include "BinnedBlinks.h"
#include "ZernikePsf.h"
#include "Illumination.h"
#include "Parallel.h"
#include "dyes.h"
#include "jsmath.h"
#include "microtubules.h"
#include "npcs.h"
#include "params.h"
#include "world.h"
#include "cells.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

extern "C" {
#include "oracle.h"
}

namespace {

sim::PsfGeneratorRequest ToRequest(const orc_psf_request* r)
{
   sim::PsfGeneratorRequest q;
   q.model = sim::PsfModelKind::GibsonLanniZernike;
   q.wavelengthNm = r->wavelengthNm;
   q.na = r->na;
   q.immersionIndex = r->immersionIndex;
   q.sampleIndex = r->sampleIndex;
   q.workingDistanceUm = r->workingDistanceUm;
   q.sampleDepthNm = r->sampleDepthNm;
   q.pixelSizeNm = r->pixelSizeNm;
   // Every digit (FormatZernikeCoefficients keeps 6 significant digits; std::stod round-trips %.17g exactly).
   std::string z;
   for (int j = 0; j < 28; ++j)
   {
      char buf[40];
      std::snprintf(buf, sizeof buf, "%s%.17g", j ? "," : "", r->zernike[j]);
      z += buf;
   }
   q.zernikeCoefficients = z;
   q.maskType = r->maskType == 1 ? sim::PsfMaskType::DoubleHelix : sim::PsfMaskType::None;
   q.maskModes = r->maskModes;
   q.maskWaist = r->maskWaist;
   q.oversampling = r->oversampling;
   q.kernelHalfWidthPx = r->kernelHalfWidthPx;
   q.nz = r->nz;
   q.zStepNm = r->zStepNm;
   q.pupilSamples = r->pupilSamples;
   q.interpMode = static_cast<sim::PsfInterpMode>(r->interpMode);
   return q;
}

sim::CameraNoiseParams ToCamera(const orc_camera* c)
{
   sim::CameraNoiseParams p;
   p.quantumEfficiency = c->quantumEfficiency;
   p.darkCurrentElectrons = c->darkCurrentElectrons;
   p.gainPhotonsPerAdu = c->gainPhotonsPerAdu;
   p.readNoiseElectrons = c->readNoiseElectrons;
   p.emccd = c->emccd != 0;
   p.emGain = c->emGain;
   p.cicElectrons = c->cicElectrons;
   p.bitDepth = c->bitDepth;
   return p;
}

// A kernel cache holding caller-given sum-1 planes (PsfKernelCache::SetData; block sums are made on first use).
sim::PsfKernelCache MakeCache(const float* planes, int nz, int size, int os, double zStepNm, int interp)
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
   d.planes.resize(static_cast<size_t>(nz));
   const size_t plane = static_cast<size_t>(size) * size;
   for (int z = 0; z < nz; ++z)
      d.planes[static_cast<size_t>(z)].assign(planes + z * plane, planes + (z + 1) * plane);
   c.SetData(std::move(d));
   c.valid = true;
   return c;
}

} // namespace

extern "C" {

// ComputeZernikePsfPlanes (Sim/ZernikePsf.cpp:150): raw planes.
int32_t ref_psf_planes(const orc_psf_request* r, float* out)
{
   std::vector<std::vector<float>> planes;
   std::string err;
   if (!sim::ComputeZernikePsfPlanes(ToRequest(r), planes, err))
      return -1;
   size_t o = 0;
   for (const auto& p : planes)
   {
      std::memcpy(out + o, p.data(), p.size() * sizeof(float));
      o += p.size();
   }
   return 0;
}

// ZernikePupilSamples (Sim/ZernikePsf.cpp:134-148): the pupil samples per axis M of a request.
int32_t ref_zernike_pupil_samples(const orc_psf_request* r)
{
   return sim::ZernikePupilSamples(ToRequest(r));
}

// BuildZernikePsfKernelCache (Sim/ZernikePsf.cpp:382): sum-1 planes and every plane's row-major block sums.
int32_t ref_psf_kernel_cache(const orc_psf_request* r, float* planesOut, float* blockSumsOut)
{
   sim::PsfKernelCache cache;
   std::string err;
   if (!sim::BuildZernikePsfKernelCache(ToRequest(r), cache, err))
      return -1;
   const size_t plane = static_cast<size_t>(cache.sizeOversampled) * cache.sizeOversampled;
   const size_t bplane = static_cast<size_t>(cache.blockSumWidth) * cache.blockSumWidth;
   for (int z = 0; z < cache.nz; ++z)
   {
      if (planesOut)
         std::memcpy(planesOut + z * plane, cache.Planes()[static_cast<size_t>(z)].data(), plane * sizeof(float));
      if (blockSumsOut)
      {
         const std::vector<float> b = cache.BlockSums(z).RowMajor();
         std::memcpy(blockSumsOut + z * bplane, b.data(), bplane * sizeof(float));
      }
   }
   return 0;
}

// BuildBlockSums (Sim/PsfGeneratorBridge.cpp:1021).
void ref_block_sums(const float* kernel, int32_t n, int32_t os, float* out)
{
   const std::vector<float> b = sim::BuildBlockSums(kernel, n, os);
   std::memcpy(out, b.data(), b.size() * sizeof(float));
}

// ApplyNoiseChain (Sim/SMLMNoise.cpp:142); maps NULL = none.
void ref_noise_chain(const float* photons, uint16_t* out, uint32_t w, uint32_t h, const orc_camera* cam,
                     const float* offset, const float* gain, const float* read, uint32_t seed, uint32_t frame,
                     uint32_t pixelBase, int32_t parallel)
{
   const size_t n = static_cast<size_t>(w) * h;
   std::vector<float> img(photons, photons + n);
   std::vector<uint16_t> adu;
   sim::PixelOffsetMap om;
   sim::PixelGainMap gm;
   sim::PixelReadNoiseMap rm;
   if (offset)
   {
      om.width = w;
      om.height = h;
      om.offset.assign(offset, offset + n);
   }
   if (gain)
   {
      gm.width = w;
      gm.height = h;
      gm.gainPhotonsPerAdu.assign(gain, gain + n);
   }
   if (read)
   {
      rm.width = w;
      rm.height = h;
      rm.readNoiseElectrons.assign(read, read + n);
   }
   sim::ApplyNoiseChain(img, adu, w, h, ToCamera(cam), om, gm, rm, seed, frame, parallel != 0, pixelBase);
   std::memcpy(out, adu.data(), n * sizeof(uint16_t));
}

// One blink event of the shim's flat event array (sim::BlinkEvent's fields used on the hot path).
struct ref_event
{
   double xUm, yUm, zNm, tStart, tEnd, brightness;
};

static std::vector<sim::BlinkEvent> ToEvents(const ref_event* ev, int32_t n)
{
   std::vector<sim::BlinkEvent> out(static_cast<size_t>(n));
   for (int32_t i = 0; i < n; ++i)
   {
      out[static_cast<size_t>(i)].xUm = ev[i].xUm;
      out[static_cast<size_t>(i)].yUm = ev[i].yUm;
      out[static_cast<size_t>(i)].zNm = ev[i].zNm;
      out[static_cast<size_t>(i)].tStart = ev[i].tStart;
      out[static_cast<size_t>(i)].tEnd = ev[i].tEnd;
      out[static_cast<size_t>(i)].brightness = ev[i].brightness;
   }
   return out;
}

// CollectFrameEmitters (Sim/SMLMSimulation.cpp:55): the frame's emitters (x, y, photons, zIndex); returns the count.
int32_t ref_collect_emitters(const ref_event* ev, int32_t nEv, int64_t frame, uint32_t w, uint32_t h,
                             double pixelSizeNm, double photonsPerBlink, int32_t nzK, double zStepNm,
                             double globalZUm, orc_emitter* out, int32_t cap)
{
   sim::PsfKernelCache c;
   c.nz = nzK;
   c.zStepNm = zStepNm;
   std::vector<sim::FrameEmitter> ems;
   sim::CollectFrameEmitters(ToEvents(ev, nEv), frame, w, h, pixelSizeNm, photonsPerBlink, 0.0, 0.0, &c, globalZUm,
                             nullptr, ems);
   const int32_t n = static_cast<int32_t>(ems.size());
   for (int32_t i = 0; i < n && i < cap; ++i)
   {
      out[i].xPx = ems[static_cast<size_t>(i)].xPx;
      out[i].yPx = ems[static_cast<size_t>(i)].yPx;
      out[i].photons = ems[static_cast<size_t>(i)].photons;
      out[i].zIndex = ems[static_cast<size_t>(i)].zIndex;
      out[i].kernel = 0;
   }
   return n;
}

// RenderPhotonImage (Sim/SMLMSimulation.cpp:103) with a kernel cache of the given sum-1 planes; halo cut by
// WithHaloCut(threshold) when threshold > 0.
int32_t ref_render_frame(float* img, uint32_t w, uint32_t h, const ref_event* ev, int32_t nEv, int64_t frame,
                         double pixelSizeNm, double photonsPerBlink, double bg, const float* planes, int32_t nz,
                         int32_t size, int32_t os, double zStepNm, int32_t interp, double globalZUm,
                         double haloThreshold, double minPhotons, int32_t parallel)
{
   sim::PsfKernelCache c = MakeCache(planes, nz, size, os, zStepNm, interp);
   if (haloThreshold > 0.0)
      c = sim::WithHaloCut(c, haloThreshold);
   std::vector<float> out;
   sim::RenderExtras x;
   x.parallel = parallel != 0;
   x.minPhotons = minPhotons;
   sim::RenderPhotonImage(out, w, h, ToEvents(ev, nEv), frame, pixelSizeNm, 1.0, photonsPerBlink, bg, 0.0, 0.0, &c,
                          globalZUm, nullptr, nullptr, &x);
   std::memcpy(img, out.data(), out.size() * sizeof(float));
   return 0;
}

// WithHaloCut (Sim/PsfGeneratorBridge.cpp:1186): the spans (nz x (2 camRad + 1) each) and radius2 per plane.
int32_t ref_halo_spans(const float* planes, int32_t nz, int32_t size, int32_t os, int32_t interp, double threshold,
                       int16_t* lo, int16_t* hi, int32_t* radius2)
{
   sim::PsfKernelCache c = sim::WithHaloCut(MakeCache(planes, nz, size, os, 100.0, interp), threshold);
   if (!c.halo)
      return -1;
   std::memcpy(lo, c.halo->lo.data(), c.halo->lo.size() * sizeof(int16_t));
   std::memcpy(hi, c.halo->hi.data(), c.halo->hi.size() * sizeof(int16_t));
   std::memcpy(radius2, c.halo->radius2.data(), c.halo->radius2.size() * sizeof(int32_t));
   return c.halo->camRad;
}

// CollectGpuEmitters (Sim/SMLMSimulation.cpp:219): 64-byte records (sim::GpuSplatEmitter); returns the count.
int32_t ref_collect_gpu_emitters(const ref_event* ev, int32_t nEv, int64_t frame, uint32_t w, uint32_t h,
                                 double pixelSizeNm, double photonsPerBlink, const float* planes, int32_t nz,
                                 int32_t size, int32_t os, double zStepNm, int32_t interp, double globalZUm,
                                 void* out, int32_t cap)
{
   sim::PsfKernelCache c = MakeCache(planes, nz, size, os, zStepNm, interp);
   std::vector<sim::GpuSplatEmitter> g;
   sim::CollectGpuEmitters(ToEvents(ev, nEv), frame, w, h, pixelSizeNm, photonsPerBlink, 0.0, 0.0, c, globalZUm,
                           nullptr, g);
   const int32_t n = static_cast<int32_t>(g.size());
   std::memcpy(out, g.data(), static_cast<size_t>(n < cap ? n : cap) * sizeof(sim::GpuSplatEmitter));
   return n;
}

// KernelWidefieldPsf (Sim/WidefieldRender.cpp:272-323): radius and grid-pitch kernel of plane p.
int32_t ref_wf_kernel(const float* planes, int32_t nz, int32_t size, int32_t os, double zStepNm, int32_t upscale,
                      int32_t p, int32_t R, float* out)
{
   sim::PsfKernelCache c = MakeCache(planes, nz, size, os, zStepNm, 0);
   sim::KernelWidefieldPsf psf(c, upscale);
   if (R <= 0)
      return psf.Radius(0, nz - 1);
   std::vector<float> k;
   psf.Kernel(p, R, k);
   std::memcpy(out, k.data(), k.size() * sizeof(float));
   return R;
}

// Frames rendered and noised as FluorescenceMovie::Render batches them (Sim/ScopeMovie.cpp:7649-7653): the frames
// of a batch in parallel, each RenderPhotonImage (serial) + ApplyNoiseChain. events: one flat list; frame f renders
// the events overlapping [f, f+1). The timed CPU baseline.
int32_t ref_render_noise_frames(const ref_event* ev, int32_t nEv, const int32_t* evStart, int32_t nFrames,
                                uint32_t w, uint32_t h, double pixelSizeNm, double bg, const float* planes,
                                int32_t nz, int32_t size, int32_t os, double zStepNm, int32_t interp,
                                const orc_camera* cam, const float* offset, uint32_t seed, uint16_t* out)
{
   sim::PsfKernelCache c = MakeCache(planes, nz, size, os, zStepNm, interp);
   c.PrepareSplat();
   const sim::CameraNoiseParams cp = ToCamera(cam);
   sim::PixelOffsetMap om;
   if (offset)
   {
      om.width = w;
      om.height = h;
      om.offset.assign(offset, offset + static_cast<size_t>(w) * h);
   }
   const sim::PixelGainMap gm;
   const sim::PixelReadNoiseMap rm;
   const size_t npx = static_cast<size_t>(w) * h;
   sim::ParallelFor(static_cast<unsigned>(nFrames), [&](unsigned f) {
      const std::vector<sim::BlinkEvent> events = ToEvents(ev + evStart[f], evStart[f + 1] - evStart[f]);
      std::vector<float> img;
      sim::RenderPhotonImage(img, w, h, events, static_cast<long>(f), pixelSizeNm, 1.0, 1.0, bg, 0.0, 0.0, &c, 0.0);
      std::vector<uint16_t> adu;
      sim::ApplyNoiseChain(img, adu, w, h, cp, om, gm, rm, seed, f, false, 0);
      std::memcpy(out + f * npx, adu.data(), npx * sizeof(uint16_t));
   });
   return 0;
}

// ComputeZernikePsfPlanes over many requests, requests in sequence and planes in parallel as the reference runs
// them (Sim/ZernikePsf.cpp:236). The timed CPU baseline of the PSF library.
int32_t ref_psf_library(const orc_psf_request* reqs, int32_t nReq, float* out)
{
   size_t o = 0;
   for (int32_t i = 0; i < nReq; ++i)
   {
      std::vector<std::vector<float>> planes;
      std::string err;
      if (!sim::ComputeZernikePsfPlanes(ToRequest(reqs + i), planes, err))
         return -1;
      for (const auto& p : planes)
      {
         if (out)
            std::memcpy(out + o, p.data(), p.size() * sizeof(float));
         o += p.size();
      }
   }
   return 0;
}

// WidefieldScene::UpdateFromGrid + FocusSeries (Sim/WidefieldRender.h:458,588) for a persistent population on a
// nx x ny x nzD grid at 100 nm x 100 nm x dzNm, flat illumination over exactly the FOV (no margin), unit dose and
// collection efficiency (the spec's mean-field convention, Sim/WidefieldRender.h:258-263), every dye plane in every
// focus (no slab), the kernel radius cap at kernelCapUm. Foci: the nF world heights focusUm. out: nF x ny x nx, the
// images of the FOV cells (WidefieldImages::persistent). Returns the FFT grid edge the scene chose (>0) or -1.
int32_t ref_wf_focus_series(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzNm,
                            const float* planes, int32_t nzP, int32_t size, int32_t os, double zStepNm,
                            double kernelCapUm, const double* focusUm, int32_t nF, float* out)
{
   sim::PsfKernelCache c = MakeCache(planes, nzP, size, os, zStepNm, 0);
   sim::KernelWidefieldPsf psf(c, 1);
   sim::WidefieldSceneSpec spec;
   spec.pixelUm = 0.1;
   spec.originXUm = 0.05;   // pixel (0, 0)'s centre: the FOV starts on a cell edge (frac 0)
   spec.originYUm = 0.05;
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
   sim::WidefieldScene sc;
   std::string err;
   if (!sc.UpdateFromGrid(G, ill, spec, psf, err))
      return -1;
   std::vector<double> foci(focusUm, focusUm + nF);
   std::vector<sim::WidefieldImages> images;
   if (!sc.FocusSeries(foci, images, err))
      return -1;
   for (int f = 0; f < nF; ++f)
   {
      const sim::WidefieldImages& im = images[static_cast<size_t>(f)];
      if (im.persistent.size() != static_cast<size_t>(nx) * ny)
         return -1;
      std::memcpy(out + static_cast<size_t>(f) * nx * ny, im.persistent.data(), im.persistent.size() * sizeof(float));
   }
   return 1;
}

// WidefieldDriftFrames (Sim/WidefieldRender.cpp:2937-2977) over a spectra-keeping scene of a persistent population
// (density as ref_wf_focus_series): Begin(DriftRange(drift), {baseFocusUm}), Refresh, then Render of each frame's
// drift (x, y, z nm: drift3[3 f ..]) into outCam (nFrames x ny x nx, from zero). Also what the frames are made from:
// the focus series on the drift's focus grid (foci base - DzNm(k) / 1000), each image's persistent spectrum
// (nImages x NY x (NX/2 + 1) complex floats, up to specCap floats) and geometry geo = {NX, NY, fovX0, fovY0, cw, ch,
// upscale, nImages, k0}, frac = {fracX, fracY}, pitchNm. Returns 0, -1 on failure, -2 when specCap is too small.
int32_t ref_wf_drift_frames(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzNm,
                            const float* planes, int32_t nzP, int32_t size, int32_t os, double zStepNm,
                            double kernelCapUm, double baseFocusUm, const double* drift3, int32_t nFrames,
                            float* outCam, float* specOut, size_t specCap, int32_t* geo, double* frac, double* pitchNm)
{
   sim::PsfKernelCache c = MakeCache(planes, nzP, size, os, zStepNm, 0);
   sim::KernelWidefieldPsf psf(c, 1);
   sim::WidefieldSceneSpec spec;
   spec.pixelUm = 0.1;
   spec.originXUm = 0.05;
   spec.originYUm = 0.05;
   spec.width = static_cast<unsigned>(nx);
   spec.height = static_cast<unsigned>(ny);
   spec.focusWorldUm = baseFocusUm;
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
   sim::WidefieldScene sc;
   sc.SetKeepSpectra(true);
   std::string err;
   if (!sc.UpdateFromGrid(G, ill, spec, psf, err))
      return -1;
   std::vector<sim::DriftNm> traj(static_cast<size_t>(nFrames));
   for (int f = 0; f < nFrames; ++f)
   {
      traj[static_cast<size_t>(f)].x = drift3[3 * f];
      traj[static_cast<size_t>(f)].y = drift3[3 * f + 1];
      traj[static_cast<size_t>(f)].z = drift3[3 * f + 2];
   }
   const sim::DriftBounds b = sim::DriftRange(traj);
   sim::WidefieldDriftFrames df;
   df.Begin(b, {baseFocusUm});
   if (!df.Refresh(sc, err))
      return -1;
   const size_t npx = static_cast<size_t>(nx) * ny;
   for (int f = 0; f < nFrames; ++f)
   {
      std::vector<float> cam(npx, 0.0f);
      df.Render(sc, 0, traj[static_cast<size_t>(f)], {}, cam);
      if (cam.size() != npx)
         return -1;
      std::memcpy(outCam + f * npx, cam.data(), npx * sizeof(float));
   }
   const sim::DriftFocusGrid grid = sim::DriftFocusGrid::For(b);
   std::vector<double> foci;
   for (int k = 0; k < grid.n; ++k)
      foci.push_back(baseFocusUm - grid.DzNm(k) / 1000.0);
   std::vector<sim::WidefieldImages> series;
   if (!sc.FocusSeries(foci, series, err) || series.empty() || !series[0].HasSpectra())
      return -1;
   const sim::WidefieldImages& i0 = series[0];
   const size_t n = i0.fft->SpecSize();
   geo[0] = static_cast<int32_t>(i0.fft->Nx());
   geo[1] = static_cast<int32_t>(i0.fft->Ny());
   geo[2] = static_cast<int32_t>(i0.fovX0);
   geo[3] = static_cast<int32_t>(i0.fovY0);
   geo[4] = static_cast<int32_t>(i0.cw);
   geo[5] = static_cast<int32_t>(i0.ch);
   geo[6] = static_cast<int32_t>(i0.upscale);
   geo[7] = grid.n;
   geo[8] = static_cast<int32_t>(grid.k0);
   frac[0] = i0.fracX;
   frac[1] = i0.fracY;
   *pitchNm = sc.Grid().pitchUm * 1000.0;
   if (2 * n * series.size() > specCap)
      return -2;
   for (size_t m = 0; m < series.size(); ++m)
   {
      if (series[m].specPersistent.size() != n)
         return -1;
      std::memcpy(specOut + 2 * n * m, series[m].specPersistent.data(), n * sizeof(sim::cfloat));
   }
   return 0;
}

// BinnedBlinkRenderer (Sim/BinnedBlinks.cpp:94-246): Setup with the caller's planes (null: the Gaussian of sigmaPx),
// cut by WithHaloCut at haloThreshold when > 0, then Render of each frame's emitters (em: n x 4 doubles xPx, yPx,
// photons, zIndex; frame f holds [start[f], start[f+1])) onto img (nFrames x H x W, added to). uR: u and R out.
int32_t ref_binned_render(const float* planes, int32_t nz, int32_t size, int32_t os, double haloThreshold,
                          double sigmaPx, int32_t W, int32_t H, int32_t upscale, const double* em, const int32_t* start,
                          int32_t nFrames, float* img, int32_t* uR)
{
   sim::PsfKernelCache c;
   if (planes)
   {
      c = MakeCache(planes, nz, size, os, 100.0, 0);
      if (haloThreshold > 0.0)
         c = sim::WithHaloCut(c, haloThreshold);
   }
   sim::BinnedBlinkRenderer r;
   if (!r.Setup(planes ? &c : nullptr, sigmaPx, static_cast<unsigned>(W), static_cast<unsigned>(H), upscale))
      return -1;
   uR[0] = r.Upscale();
   uR[1] = r.RadiusCells();
   const size_t npx = static_cast<size_t>(W) * H;
   for (int f = 0; f < nFrames; ++f)
   {
      std::vector<sim::FrameEmitter> ems;
      for (int i = start[f]; i < start[f + 1]; ++i)
      {
         sim::FrameEmitter e{};
         e.xPx = em[4 * i];
         e.yPx = em[4 * i + 1];
         e.photons = em[4 * i + 2];
         e.overlap = 1.0;
         e.zIndex = static_cast<int>(em[4 * i + 3]);
         ems.push_back(e);
      }
      std::vector<float> frame(img + f * npx, img + (f + 1) * npx);
      r.Render(ems, frame);
      std::memcpy(img + f * npx, frame.data(), npx * sizeof(float));
   }
   return 0;
}

// DyeH1 / NpcH1 / LaminH1 (core/src/dyes.cpp:175-222): which 0, 1, 2.
uint32_t ref_dye_h1(int32_t which, uint32_t seed, int32_t cx, int32_t cy, int32_t mtIndex)
{
   return which == 0 ? isc::DyeH1(seed, cx, cy, mtIndex) : which == 1 ? isc::NpcH1(seed, cx, cy)
                                                                     : isc::LaminH1(seed, cx, cy);
}

// jsm::sin / cos / exp / log (core/src/jsmath_fdlibm.cpp) of n inputs: fn 0 sin, 1 cos, 2 exp, 3 log.
void ref_jsm(int32_t fn, const double* x, int64_t n, double* out)
{
   for (int64_t i = 0; i < n; ++i)
      out[i] = fn == 0 ? isc::jsm::sin(x[i]) : fn == 1 ? isc::jsm::cos(x[i]) : fn == 2 ? isc::jsm::exp(x[i])
                                                                                  : isc::jsm::log(x[i]);
}

static isc::Label MakeLabel(int32_t mode, int32_t preState, const orc_kinetics* kin)
{
   isc::Label l;
   l.mode = static_cast<isc::LabelMode>(mode);
   l.preState = preState != 0;
   l.kin.activationRatePerSec = kin->activationRatePerSec;
   l.kin.onSec = kin->onSec;
   l.kin.offSec = kin->offSec;
   l.kin.bleachProb = kin->bleachProb;
   l.kin.photonCV = kin->photonCV;
   l.kin.initialOnSec = kin->initialOnSec;
   return l;
}

static isc::Kinetics KinOf(const orc_kinetics& k)
{
   isc::Kinetics q;
   q.activationRatePerSec = k.activationRatePerSec;
   q.onSec = k.onSec;
   q.offSec = k.offSec;
   q.bleachProb = k.bleachProb;
   q.photonCV = k.photonCV;
   q.initialOnSec = k.initialOnSec;
   return q;
}

// A KineticsHistory of nSeg segments (tStart[i], segKin[i]); nullptr for nSeg 0 (none).
static const isc::KineticsHistory* HistOf(int32_t nSeg, const double* tStart, const orc_kinetics* segKin,
                                          isc::KineticsHistory& h)
{
   if (nSeg <= 0)
      return nullptr;
   h.tStart.assign(tStart, tStart + nSeg);
   h.kin.clear();
   for (int32_t i = 0; i < nSeg; ++i)
      h.kin.push_back(KinOf(segKin[i]));
   return &h;
}

// LabelSchedule (core/src/dyes.cpp:440) for nDyes addresses (h1, k, n as 3 x uint32), blinks only, under the history
// of nSeg segments (0: none): counts[i] per dye, out = every blink (tOn, tOff, brightness) in dye order, up to cap.
// Returns the total.
int64_t ref_label_blinks(const uint32_t* dyes, int32_t nDyes, int32_t mode, const orc_kinetics* kin, double tMax,
                         int32_t* counts, double* out, int64_t cap, int32_t nSeg, const double* tStart,
                         const orc_kinetics* segKin)
{
   const isc::Label l = MakeLabel(mode, 0, kin);
   isc::KineticsHistory hh;
   const isc::KineticsHistory* hist = HistOf(nSeg, tStart, segKin, hh);
   std::vector<isc::Blink> b;
   int64_t total = 0;
   for (int32_t i = 0; i < nDyes; ++i)
   {
      b.clear();
      isc::LabelSchedule(dyes[3 * i], (int32_t)dyes[3 * i + 1], (int32_t)dyes[3 * i + 2], l, &b, nullptr, tMax,
                         hist);
      counts[i] = (int32_t)b.size();
      for (const isc::Blink& x : b)
      {
         if (total < cap)
         {
            out[3 * total] = x.tOn;
            out[3 * total + 1] = x.tOff;
            out[3 * total + 2] = x.brightness;
         }
         ++total;
      }
   }
   return total;
}

// LabelSchedule's continuous windows for nDyes addresses (history as ref_label_blinks'): has[i] windows (0 or 1),
// w = 4 doubles per dye.
void ref_label_cont(const uint32_t* dyes, int32_t nDyes, int32_t mode, int32_t preState, const orc_kinetics* kin,
                    int32_t* has, double* w, int32_t nSeg, const double* tStart, const orc_kinetics* segKin)
{
   const isc::Label l = MakeLabel(mode, preState, kin);
   isc::KineticsHistory hh;
   const isc::KineticsHistory* hist = HistOf(nSeg, tStart, segKin, hh);
   std::vector<isc::ContWindow> c;
   for (int32_t i = 0; i < nDyes; ++i)
   {
      c.clear();
      isc::LabelSchedule(dyes[3 * i], (int32_t)dyes[3 * i + 1], (int32_t)dyes[3 * i + 2], l, nullptr, &c, INFINITY,
                         hist);
      has[i] = (int32_t)c.size();
      if (!c.empty())
      {
         w[4 * i] = c[0].tOn;
         w[4 * i + 1] = c[0].tOff;
         w[4 * i + 2] = c[0].state;
         w[4 * i + 3] = c[0].aux;
      }
   }
}

// PersistentBlinks (bins 0: window t0, t1) or PersistentBlinksInBins (bins [binLo, binHi]) (core/src/dyes.cpp:636,
// :648) for nDyes addresses: counts per dye, out = 5 doubles per blink (tOn, tOff, brightness, bin, j; bin and j 0
// for the window form) up to cap, under the history of nSeg segments (0: none). Returns the total.
int64_t ref_persistent_blinks(const uint32_t* dyes, int32_t nDyes, const orc_kinetics* kin, int32_t bins, double t0,
                              double t1, int64_t binLo, int64_t binHi, int32_t* counts, double* out, int64_t cap,
                              int32_t nSeg, const double* tStart, const orc_kinetics* segKin)
{
   const isc::Label l = MakeLabel(2, 0, kin);
   isc::KineticsHistory hh;
   const isc::KineticsHistory* hist = HistOf(nSeg, tStart, segKin, hh);
   int64_t total = 0;
   for (int32_t i = 0; i < nDyes; ++i)
   {
      const uint32_t h1 = dyes[3 * i];
      const int32_t k = (int32_t)dyes[3 * i + 1], n = (int32_t)dyes[3 * i + 2];
      std::vector<double> rec;
      if (bins)
      {
         std::vector<isc::BinBlink> b;
         isc::PersistentBlinksInBins(h1, k, n, l.kin, (long)binLo, (long)binHi, b, hist);
         for (const isc::BinBlink& x : b)
            rec.insert(rec.end(), { x.tOn, x.tOff, x.brightness, (double)x.bin, (double)x.j });
      }
      else
      {
         std::vector<isc::Blink> b;
         isc::PersistentBlinks(h1, k, n, l.kin, t0, t1, b, hist);
         for (const isc::Blink& x : b)
            rec.insert(rec.end(), { x.tOn, x.tOff, x.brightness, 0.0, 0.0 });
      }
      counts[i] = (int32_t)(rec.size() / 5);
      for (size_t r = 0; r < rec.size(); r += 5, ++total)
         if (total < cap)
            std::memcpy(out + 5 * total, rec.data() + r, 5 * sizeof(double));
   }
   return total;
}

} // extern "C"

namespace {

// A WidefieldAccelerator that records each job the scene sends (in the byte layout documented at
// ref_wf_capture_jobs) and answers with zero images; planes count as resident once a job has carried them.
class CaptureHost : public sim::WidefieldAccelerator
{
public:
   bool HasPlane(unsigned long long geometry, unsigned long long key) const override
   {
      return geometry == geometry_ && keys_.count(key) != 0;
   }
   bool Images(const sim::WidefieldGpuJob& job, std::vector<std::vector<float>>& out, std::string&) override
   {
      if (job.geometry != geometry_)
      {
         keys_.clear();
         geometry_ = job.geometry;
      }
      const unsigned W = job.NX / 2 + 1;
      for (uint32_t v : {job.NX, job.NY, job.nx, job.ny, job.fovX0, job.fovY0, job.cw, job.ch})
         Put(v);
      Put(job.fracX);
      Put(job.fracY);
      Put(static_cast<uint64_t>(job.geometry));
      Put(static_cast<int32_t>(job.kernels.size()));
      Put(static_cast<int32_t>(job.planes.size()));
      Put(static_cast<int32_t>(job.channels.size()));
      Put(static_cast<int32_t>(job.hasPersistent));
      for (const auto& k : job.kernels)
      {
         Put(static_cast<int32_t>(k.p));
         if (k.spec->size() != static_cast<size_t>(W) * job.NY)
            return false;
         Bytes(k.spec->data(), k.spec->size() * sizeof(sim::cfloat));
      }
      for (const auto& p : job.planes)
      {
         Put(static_cast<uint64_t>(p.key));
         Put(static_cast<uint32_t>(p.cells.size()));
         Bytes(p.cells.data(), p.cells.size() * sizeof(uint32_t));
         Bytes(p.values.data(), p.values.size() * sizeof(float));
         keys_.insert(p.key);
      }
      for (const auto& ch : job.channels)
      {
         Put(static_cast<uint32_t>(ch.size()));
         for (const auto& d : ch)
         {
            Put(static_cast<uint64_t>(d.key));
            Put(static_cast<int32_t>(d.p0));
            Put(static_cast<int32_t>(d.p1));
            Put(d.w0);
            Put(d.w1);
            Put(static_cast<int32_t>(d.two));
         }
      }
      out.assign(job.channels.size(), std::vector<float>(static_cast<size_t>(job.cw) * job.ch, 0.0f));
      ++jobs;
      return true;
   }
   std::vector<uint8_t> blob;
   int jobs = 0;

private:
   template <typename T>
   void Put(T v)
   {
      Bytes(&v, sizeof v);
   }
   void Bytes(const void* p, size_t n)
   {
      const uint8_t* b = static_cast<const uint8_t*>(p);
      blob.insert(blob.end(), b, b + n);
   }
   unsigned long long geometry_ = ~0ull;
   std::set<unsigned long long> keys_;
};

} // namespace

extern "C" {

// The jobs a GPU-mode WidefieldScene sends its accelerator for FocusSeries over the foci (one per focus, the
// planes resident after the first), and the same GPU-mode scene's own CPU images: a persistent population
// (density [nzD][ny][nx]) under flat illumination, the FOV's origin moved by (dxUm, dyUm) off the cell edges.
// Job layout (little-endian, packed): u32 NX NY nx ny fovX0 fovY0 cw ch; f64 fracX fracY; u64 geometry;
// i32 nKernels nPlanes nChannels hasPersistent; per kernel i32 p + (NX/2+1) NY complex f32; per plane u64 key,
// u32 n, u32 cells[n], f32 values[n]; per channel u32 nDeps, per dep u64 key, i32 p0 p1, f32 w0 w1, i32 two.
// Returns the number of jobs (-1 on failure, -2 when blobCap is too small); *blobUsed = bytes written.
int32_t ref_wf_capture_jobs(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzNm,
                            const float* planes, int32_t nzP, int32_t size, int32_t os, double zStepNm,
                            double kernelCapUm, const double* focusUm, int32_t nF, double dxUm, double dyUm,
                            uint8_t* blob, size_t blobCap, size_t* blobUsed, float* outCpu)
{
   sim::PsfKernelCache c = MakeCache(planes, nzP, size, os, zStepNm, 0);
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
   CaptureHost host;
   sim::WidefieldScene gs, cs;
   gs.SetGpuMode(true);
   cs.SetGpuMode(true);
   gs.SetAccelerator(&host);
   std::string err;
   if (!gs.UpdateFromGrid(G, ill, spec, psf, err) || !cs.UpdateFromGrid(G, ill, spec, psf, err))
      return -1;
   std::vector<double> foci(focusUm, focusUm + nF);
   std::vector<sim::WidefieldImages> gi, ci;
   if (!gs.FocusSeries(foci, gi, err) || !cs.FocusSeries(foci, ci, err) || !gs.UsingAccelerator())
      return -1;
   for (int f = 0; f < nF; ++f)
   {
      const sim::WidefieldImages& im = ci[static_cast<size_t>(f)];
      if (im.persistent.size() != static_cast<size_t>(nx) * ny)
         return -1;
      std::memcpy(outCpu + static_cast<size_t>(f) * nx * ny, im.persistent.data(), im.persistent.size() * sizeof(float));
   }
   *blobUsed = host.blob.size();
   if (host.blob.size() > blobCap)
      return -2;
   std::memcpy(blob, host.blob.data(), host.blob.size());
   return host.jobs;
}

} // extern "C"

extern "C" {

// WidefieldImages::Render (Sim/WidefieldRender.cpp:547-580) of channel images (channels: [nCh][ch][cw], channel 0
// the persistent one when hasPersistent) with bleach coefficients a (nA doubles), added to cam ((ch/u) x (cw/u)).
int32_t ref_wf_images_render(const float* channels, int32_t nCh, int32_t hasPersistent, uint32_t cw, uint32_t ch,
                             int32_t upscale, const double* a, int32_t nA, float* cam)
{
   sim::WidefieldImages im;
   im.cw = cw;
   im.ch = ch;
   im.upscale = static_cast<unsigned>(upscale);
   const size_t n = static_cast<size_t>(cw) * ch;
   int c = 0;
   if (hasPersistent)
   {
      im.persistent.assign(channels, channels + n);
      c = 1;
   }
   for (; c < nCh; ++c)
      im.bleach.emplace_back(channels + static_cast<size_t>(c) * n, channels + static_cast<size_t>(c + 1) * n);
   std::vector<float> out(cam, cam + static_cast<size_t>(cw / std::max(1, upscale)) * (ch / std::max(1, upscale)));
   im.Render(std::vector<double>(a, a + nA), out);
   std::memcpy(cam, out.data(), out.size() * sizeof(float));
   return 0;
}

} // extern "C"

extern "C" {

// NoiseSetup's maps (Sim/ScopeMovie.cpp:973-996, an anonymous struct there): the three Generate calls on one
// std::mt19937_64(seed) in its order, whole sensor (sensorH 0) or a window.
int32_t ref_noise_setup(int64_t seed, uint32_t W, uint32_t H, uint32_t sensorH, uint32_t rowY0, double offsetAdu,
                        double offsetStd, double gain, double gainStd, double read, double readStd, float* off,
                        float* gn, float* rn)
{
   std::mt19937_64 rng(static_cast<uint64_t>(seed));
   sim::PixelOffsetMap o;
   sim::PixelGainMap g;
   sim::PixelReadNoiseMap r;
   if (sensorH)
   {
      o.Generate(W, sensorH, offsetAdu, offsetStd, rng, rowY0, H);
      g.Generate(W, sensorH, gain, gainStd, rng, rowY0, H);
      r.Generate(W, sensorH, read, readStd, rng, rowY0, H);
   }
   else
   {
      o.Generate(W, H, offsetAdu, offsetStd, rng);
      g.Generate(W, H, gain, gainStd, rng);
      r.Generate(W, H, read, readStd, rng);
   }
   std::memcpy(off, o.offset.data(), o.offset.size() * sizeof(float));
   std::memcpy(gn, g.gainPhotonsPerAdu.data(), g.gainPhotonsPerAdu.size() * sizeof(float));
   std::memcpy(rn, r.readNoiseElectrons.data(), r.readNoiseElectrons.size() * sizeof(float));
   return 0;
}

int32_t ref_psf_half_width_px(double halfWidthNm, double pixelSizeNm, double wavelengthNm, double na)
{
   return sim::PsfKernelHalfWidthPx(halfWidthNm, pixelSizeNm, wavelengthNm, na);
}

int32_t ref_capped_psf_oversampling(int32_t os, double pixelSizeNm, double maxSampleNm)
{
   return sim::CappedPsfOversampling(os, pixelSizeNm, maxSampleNm);
}

} // extern "C"

extern "C" {

// ParseZernikeCoefficients (Sim/SMLMZernike.cpp:28): 1 and the 28 values, or 0 (out all zero).
int32_t ref_parse_zernike(const char* text, double* out)
{
   bool ok = false;
   const sim::ZernikeCoefficients z = sim::ParseZernikeCoefficients(text ? text : "", ok);
   std::memcpy(out, z.data(), z.size() * sizeof(double));
   return ok ? 1 : 0;
}

// FormatZernikeCoefficients (Sim/SMLMZernike.cpp:16): the text (NUL-terminated in buf); returns its length.
int32_t ref_format_zernike(const double* in, char separator, char* buf, int32_t cap)
{
   sim::ZernikeCoefficients z{};
   std::memcpy(z.data(), in, z.size() * sizeof(double));
   const std::string s = sim::FormatZernikeCoefficients(z, separator);
   if (buf && cap > 0)
   {
      const size_t n = std::min(s.size(), static_cast<size_t>(cap) - 1);
      std::memcpy(buf, s.data(), n);
      buf[n] = 0;
   }
   return static_cast<int32_t>(s.size());
}

// ComputePsfKernelCache (Sim/PsfGeneratorBridge.cpp:895) of a GibsonLanniZernike request whose coefficient string
// is zernikeText (NULL: the request's values at %.17g), with the disk store in diskDir ("" none) and memoEntries
// (0: unchanged); the log messages, one per line, into log; the sum-1 planes into planesOut (may be NULL) and the
// cache's serial into serial. Returns 0, or -1 with the error in log.
int32_t ref_compute_psf_kernel_cache(const orc_psf_request* r, const char* zernikeText, const char* diskDir,
                                     int32_t memoEntries, char* log, int32_t logCap, float* planesOut,
                                     uint64_t* serial)
{
   sim::PsfGeneratorRequest q = ToRequest(r);
   if (zernikeText)
      q.zernikeCoefficients = zernikeText;
   sim::SetPsfKernelDiskCacheDir(diskDir ? diskDir : "");
   if (memoEntries > 0)
      sim::SetPsfKernelMemoEntries(static_cast<size_t>(memoEntries));
   std::string text;
   sim::PsfKernelCache cache;
   std::string err;
   const bool ok = sim::ComputePsfKernelCache(q, cache, err, [&](const std::string& m) { text += m + "\n"; });
   if (!ok)
      text += err + "\n";
   if (log && logCap > 0)
   {
      const size_t n = std::min(text.size(), static_cast<size_t>(logCap) - 1);
      std::memcpy(log, text.data(), n);
      log[n] = 0;
   }
   if (!ok)
      return -1;
   if (planesOut)
   {
      const size_t plane = static_cast<size_t>(cache.sizeOversampled) * cache.sizeOversampled;
      for (int z = 0; z < cache.nz; ++z)
         std::memcpy(planesOut + z * plane, cache.Planes()[static_cast<size_t>(z)].data(), plane * sizeof(float));
   }
   if (serial)
      *serial = cache.Serial();
   return 0;
}

} // extern "C"

extern "C" {

// BucketEventsByFrame (Sim/SMLMSimulation.cpp:203): per frame the indices of the events it overlaps, in event order;
// counts[nFrames], then the lists concatenated into idx (cap entries). Returns the total, or -1 if over cap.
int64_t ref_bucket_events(const ref_event* ev, int32_t nEv, int64_t nFrames, int64_t* counts, uint32_t* idx,
                          int64_t cap)
{
   const std::vector<std::vector<uint32_t>> b = sim::BucketEventsByFrame(ToEvents(ev, nEv), static_cast<long>(nFrames));
   int64_t total = 0;
   for (size_t f = 0; f < b.size(); ++f)
   {
      counts[f] = static_cast<int64_t>(b[f].size());
      if (total + counts[f] > cap)
         return -1;
      std::memcpy(idx + total, b[f].data(), b[f].size() * sizeof(uint32_t));
      total += counts[f];
   }
   return total;
}

} // extern "C"

#include "tiff_writer.h"

extern "C" {

// TiffWriter (cli/tiff_writer.h), one call per method.
void* ref_tiff_open(const char* path, uint64_t expectedBytes, int32_t big)
{
   return new TiffWriter(path, expectedBytes, big != 0);
}

int32_t ref_tiff_page_u16(void* t, const uint16_t* px, uint32_t w, uint32_t h, const char* desc)
{
   return static_cast<TiffWriter*>(t)->Page(std::vector<uint16_t>(px, px + static_cast<size_t>(w) * h), w, h,
                                            desc ? desc : "") ? 0 : -1;
}

int32_t ref_tiff_page_f32(void* t, const float* px, uint32_t w, uint32_t h, const char* desc)
{
   return static_cast<TiffWriter*>(t)->Page(std::vector<float>(px, px + static_cast<size_t>(w) * h), w, h,
                                            desc ? desc : "") ? 0 : -1;
}

int32_t ref_tiff_page_rows_u16(void* t, const uint16_t* px, uint32_t rows, uint32_t w, uint32_t h, const char* desc)
{
   return static_cast<TiffWriter*>(t)->PageRows(px, rows, w, h, desc ? desc : "") ? 0 : -1;
}

int32_t ref_tiff_rows_u16(void* t, uint64_t page, uint32_t y0, uint32_t rows, const uint16_t* px)
{
   return static_cast<TiffWriter*>(t)->Rows(page, y0, rows, px) ? 0 : -1;
}

// Closes and frees; 0 when the file is whole.
int32_t ref_tiff_close(void* t)
{
   TiffWriter* w = static_cast<TiffWriter*>(t);
   const bool ok = w->Close();
   delete w;
   return ok ? 0 : -1;
}

uint64_t ref_tiff_bytes(uint64_t w, uint64_t h, uint64_t pages, uint64_t bps, uint64_t desc)
{
   return TiffWriter::Bytes(w, h, pages, bps, desc);
}

} // extern "C"

extern "C" {

// CollectFrameEmitters (Sim/SMLMSimulation.cpp:55) with every input: events as 7 doubles (x, y, z, tStart, tEnd,
// brightness, excite), drift, an illumination field (NULL: none), a kernel z geometry (nz 0: no kernel); records
// (x, y, photons, zIndex) and the clamped / total z counts. Returns the count.
int32_t ref_collect_emitters_full(const double* ev7, int32_t nEv, int64_t frame, uint32_t w, uint32_t h,
                                  double pixelSizeNm, double photonsPerBlink, double driftX, double driftY,
                                  int32_t nzK, double zStepNm, double globalZUm, const float* illum, orc_emitter* out,
                                  int32_t cap, int64_t* zCounts)
{
   std::vector<sim::BlinkEvent> evs(static_cast<size_t>(nEv));
   for (int32_t i = 0; i < nEv; ++i)
   {
      sim::BlinkEvent& e = evs[static_cast<size_t>(i)];
      const double* p = ev7 + 7 * static_cast<size_t>(i);
      e.xUm = p[0];
      e.yUm = p[1];
      e.zNm = p[2];
      e.tStart = p[3];
      e.tEnd = p[4];
      e.brightness = p[5];
      e.excite = p[6];
   }
   sim::PsfKernelCache c;
   c.nz = nzK;
   c.zStepNm = zStepNm;
   std::vector<float> field;
   if (illum)
      field.assign(illum, illum + static_cast<size_t>(w) * h);
   std::vector<sim::FrameEmitter> ems;
   long clamped = 0, total = 0;
   sim::CollectFrameEmitters(evs, frame, w, h, pixelSizeNm, photonsPerBlink, driftX, driftY, nzK > 0 ? &c : nullptr,
                             globalZUm, illum ? &field : nullptr, ems, &clamped, &total);
   const int32_t n = static_cast<int32_t>(ems.size());
   for (int32_t i = 0; i < n && i < cap; ++i)
   {
      out[i].xPx = ems[static_cast<size_t>(i)].xPx;
      out[i].yPx = ems[static_cast<size_t>(i)].yPx;
      out[i].photons = ems[static_cast<size_t>(i)].photons;
      out[i].zIndex = ems[static_cast<size_t>(i)].zIndex;
      out[i].kernel = 0;
   }
   zCounts[0] = clamped;
   zCounts[1] = total;
   return n;
}

// ---- the world's dye blocks and queries (core/src/world.{h,cpp}), for the checks of the device world ----

// A reference World of seed `seed` with the default params (core/src/params.h) and n name = value overrides
// (isc::SetParam). nullptr on an unknown name.
void* ref_world_new(uint32_t seed, const char* const* names, const double* values, int32_t n)
{
   isc::Params p;
   for (int32_t i = 0; i < n; ++i)
      if (!isc::SetParam(p, names[i], values[i]))
         return nullptr;
   return new isc::World(seed, p);
}

void ref_world_free(void* w) { delete static_cast<isc::World*>(w); }

// World::SetLabel from ISC_LABEL_COUNT (20) doubles in the C ABI's layout (core/include/insiliscope/insiliscope.h:
// 218-278): density, fluorescent fraction, mode, rate, on, off, bleach, cv, initial on, pre state, orientation mode,
// polar, azimuth, wobble, motion, off-target count, binder, flex, linker min, linker max (any NaN: the structure's
// linkage). 0, or -1 when the world refuses it.
int32_t ref_world_set_label(void* wp, int32_t s, const double* v)
{
   isc::Label l;
   l.density = v[0];
   l.fluorescentFraction = v[1];
   l.mode = static_cast<isc::LabelMode>(static_cast<int>(v[2]));
   l.kin.activationRatePerSec = v[3];
   l.kin.onSec = v[4];
   l.kin.offSec = v[5];
   l.kin.bleachProb = v[6];
   l.kin.photonCV = v[7];
   l.kin.initialOnSec = v[8];
   l.preState = v[9] != 0;
   l.orientation.mode = static_cast<isc::OrientationMode>(static_cast<int>(v[10]));
   l.orientation.polarDeg = v[11];
   l.orientation.azimuthDeg = v[12];
   l.orientation.wobbleDeg = v[13];
   l.motion = static_cast<int>(v[14]);
   l.offTargetCount = static_cast<int>(v[15]);
   l.hasLinkage = !(std::isnan(v[16]) || std::isnan(v[17]) || std::isnan(v[18]) || std::isnan(v[19]));
   if (l.hasLinkage)
      l.linkage = isc::Linkage{v[16], v[17], v[18], v[19]};
   return static_cast<isc::World*>(wp)->SetLabel(s, l) ? 0 : -1;
}

// World::SetKineticsHistory (core/src/world.cpp:933): n segment starts, kin[i * 4 + s] structure s's kinetics in
// segment i. 0, or -1 when the world refuses it.
int32_t ref_world_set_history(void* wp, const double* tStart, const orc_kinetics* kin, int32_t n)
{
   std::vector<double> t(tStart, tStart + n);
   std::vector<isc::Kinetics> k;
   for (int32_t i = 0; i < 4 * n; ++i)
      k.push_back(KinOf(kin[i]));
   return static_cast<isc::World*>(wp)->SetKineticsHistory(t, k) ? 0 : -1;
}

// The label the world keeps for structure s, linkage normalised: density, fraction, mode, pre state, rate, on, off,
// bleach, cv, initial on, binder, flex, linker min, linker max (14 doubles).
void ref_world_get_label(void* wp, int32_t s, double* out)
{
   const isc::Label& l = static_cast<isc::World*>(wp)->GetLabel(s);
   const double v[14] = {l.density, l.fluorescentFraction, (double)(int)l.mode, l.preState ? 1.0 : 0.0,
                         l.kin.activationRatePerSec, l.kin.onSec, l.kin.offSec, l.kin.bleachProb, l.kin.photonCV,
                         l.kin.initialOnSec, l.linkage.binderNm, l.linkage.binderFlexNm, l.linkage.linkerMinNm,
                         l.linkage.linkerMaxNm};
   std::memcpy(out, v, sizeof v);
}

// World::CellsInRect (its order): rows cx, cy, x, y, packRot, rOuter. Returns the count (rows past cap not written).
int32_t ref_world_cells(void* wp, double x0, double y0, double x1, double y1, double* out, int32_t cap)
{
   std::vector<isc::Cell> cells;
   static_cast<isc::World*>(wp)->CellsInRect(x0, y0, x1, y1, cells);
   for (size_t i = 0; i < cells.size() && (int32_t)i < cap; ++i)
   {
      const isc::Cell& c = cells[i];
      const double r[6] = {(double)c.cx, (double)c.cy, c.x, c.y, c.packRot, c.rOuter};
      std::memcpy(out + 6 * i, r, sizeof r);
   }
   return (int32_t)cells.size();
}

// The microtubules of the cell at chunk (cx, cy) (World::Assets): every point (x, y, z cell-local) up to capPts and
// each microtubule's point count up to capMts; *nMts = the count. Returns the total point count, -1 without a cell.
int64_t ref_world_cell_mts(void* wp, int32_t cx, int32_t cy, double* xyz, int64_t capPts, int32_t* lens,
                           int32_t capMts, int32_t* nMts)
{
   isc::World& w = *static_cast<isc::World*>(wp);
   isc::Cell c;
   if (!w.FindCell(cx, cy, c))
      return -1;
   const isc::CellAssets& A = w.Assets(c);
   int64_t total = 0;
   for (size_t i = 0; i < A.mts.size(); ++i)
   {
      const std::vector<isc::Pt3>& pts = A.mts[i].pts;
      if ((int32_t)i < capMts)
         lens[i] = (int32_t)pts.size();
      for (const isc::Pt3& q : pts)
      {
         if (total < capPts)
         {
            xyz[3 * total] = q.x;
            xyz[3 * total + 1] = q.y;
            xyz[3 * total + 2] = q.z;
         }
         ++total;
      }
   }
   *nMts = (int32_t)A.mts.size();
   return total;
}

// BuildMtFrames (core/src/microtubules.cpp:701) of n points: cum (n), then T, U, V (3 x (n - 1) each).
void ref_mt_frames(const double* xyz, int32_t n, double* cum, double* T, double* U, double* V)
{
   std::vector<isc::Pt3> pts((size_t)n);
   for (int32_t i = 0; i < n; ++i)
      pts[(size_t)i] = isc::Pt3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
   const isc::MtFrames f = isc::BuildMtFrames(pts);
   for (size_t i = 0; i < f.cum.size(); ++i)
      cum[i] = f.cum[i];
   for (size_t i = 0; i < f.T.size(); ++i)
   {
      const isc::Pt3* v[3] = {&f.T[i], &f.U[i], &f.V[i]};
      double* o[3] = {T, U, V};
      for (int q = 0; q < 3; ++q)
      {
         o[q][3 * i] = v[q]->x;
         o[q][3 * i + 1] = v[q]->y;
         o[q][3 * i + 2] = v[q]->z;
      }
   }
}

// The pores (which 1: World::CellNpcs; rows c, n, e1, e2, cand: 13 doubles) or lamina points (which 2:
// World::CellLamin; rows c, n, cand: 7 doubles) of the cell at chunk (cx, cy), up to cap rows. Returns the count, -1
// without a cell.
int32_t ref_world_cell_envelope(void* wp, int32_t which, int32_t cx, int32_t cy, double* out, int32_t cap)
{
   isc::World& w = *static_cast<isc::World*>(wp);
   if (which == 1)
   {
      std::vector<isc::Npc> v;
      if (!w.CellNpcs(cx, cy, v))
         return -1;
      for (size_t i = 0; i < v.size() && (int32_t)i < cap; ++i)
      {
         const isc::Npc& q = v[i];
         const double r[13] = {q.c.x,  q.c.y,  q.c.z,  q.n.x,  q.n.y,  q.n.z,         q.e1.x,
                               q.e1.y, q.e1.z, q.e2.x, q.e2.y, q.e2.z, (double)q.cand};
         std::memcpy(out + 13 * i, r, sizeof r);
      }
      return (int32_t)v.size();
   }
   std::vector<isc::LaminPoint> v;
   if (!w.CellLamin(cx, cy, v))
      return -1;
   for (size_t i = 0; i < v.size() && (int32_t)i < cap; ++i)
   {
      const isc::LaminPoint& q = v[i];
      const double r[7] = {q.c.x, q.c.y, q.c.z, q.n.x, q.n.y, q.n.z, (double)q.cand};
      std::memcpy(out + 7 * i, r, sizeof r);
   }
   return (int32_t)v.size();
}

// NpcAttachmentPoints (64 x 3 nm) and NpcSiteReachUm of the world's params.
void ref_world_npc_sites(void* wp, double* sites, double* reachUm)
{
   const isc::Params& p = static_cast<isc::World*>(wp)->GetParams();
   isc::NpcAttachmentPoints(p, sites);
   *reachUm = isc::NpcSiteReachUm(p);
}

// World::SitesInWindow: rows x, y, z, id, cx, cy, mtIndex, k, n, structure (10 doubles). Returns the count.
int64_t ref_world_sites(void* wp, double x0, double y0, double x1, double y1, double zMin, double zMax, double* out,
                        int64_t cap)
{
   std::vector<isc::WorldDye> v;
   static_cast<isc::World*>(wp)->SitesInWindow(x0, y0, x1, y1, zMin, zMax, v);
   for (size_t i = 0; i < v.size() && (int64_t)i < cap; ++i)
   {
      const isc::WorldDye& d = v[i];
      const double r[10] = {d.x, d.y, d.z, (double)d.id, (double)d.cx, (double)d.cy, (double)d.mtIndex,
                            (double)d.k, (double)d.n, (double)d.structure};
      std::memcpy(out + 10 * i, r, sizeof r);
   }
   return (int64_t)v.size();
}

// World::EventsInWindow (which 0; t0, t1) or World::ContinuousInWindow (which 1; tMin = t0): rows x, y, z, tOn, tOff,
// brightness, aux, id, structure, state (10 doubles). Returns the count.
int64_t ref_world_events(void* wp, int32_t which, double x0, double y0, double x1, double y1, double zMin,
                         double zMax, double t0, double t1, double* out, int64_t cap)
{
   std::vector<isc::WorldEvent> v;
   isc::World& w = *static_cast<isc::World*>(wp);
   if (which == 0)
      w.EventsInWindow(x0, y0, x1, y1, zMin, zMax, t0, t1, v);
   else
      w.ContinuousInWindow(x0, y0, x1, y1, zMin, zMax, t0, v);
   for (size_t i = 0; i < v.size() && (int64_t)i < cap; ++i)
   {
      const isc::WorldEvent& e = v[i];
      const double r[10] = {e.x, e.y, e.z, e.tOn, e.tOff, e.brightness, e.aux, (double)e.id, (double)e.structure,
                            (double)e.state};
      std::memcpy(out + 10 * i, r, sizeof r);
   }
   return (int64_t)v.size();
}

// World::Density3dInWindow into out[nx * ny * nz]; returns the total.
int64_t ref_world_density3d(void* wp, double x0, double y0, double x1, double y1, double zMin, double zMax,
                            int32_t nx, int32_t ny, int32_t nz, uint32_t mask, float* out)
{
   return static_cast<isc::World*>(wp)->Density3dInWindow(x0, y0, x1, y1, zMin, zMax, nx, ny, nz, mask, out);
}

// DyesInBlock (core/src/dyes.cpp:127) for one microtubule given as points: rows x, y, z, k, n, id, theta
// (7 doubles), up to cap. Returns the count.
int32_t ref_dyes_in_block(uint32_t seed, int32_t cx, int32_t cy, int32_t mtIndex, const double* xyz, int32_t nPts,
                          int32_t block, double density, double fluorescentFraction, const double* linkage4,
                          double* out, int32_t cap)
{
   std::vector<isc::Pt3> pts((size_t)nPts);
   for (int32_t i = 0; i < nPts; ++i)
      pts[(size_t)i] = isc::Pt3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
   const isc::MtFrames fr = isc::BuildMtFrames(pts);
   std::vector<isc::Dye> v;
   isc::DyesInBlock(seed, cx, cy, mtIndex, pts, fr, block, density, fluorescentFraction, v,
                    isc::Linkage{linkage4[0], linkage4[1], linkage4[2], linkage4[3]});
   for (size_t i = 0; i < v.size() && (int32_t)i < cap; ++i)
   {
      const isc::Dye& d = v[i];
      const double r[7] = {d.pos.x, d.pos.y, d.pos.z, (double)d.k, (double)d.n, (double)d.id, d.theta};
      std::memcpy(out + 7 * i, r, sizeof r);
   }
   return (int32_t)v.size();
}

// jsm::cbrt and the 2- and 3-argument jsm::hypot (core/src/jsmath.h) of n inputs: fn 0 cbrt(x), 1 hypot(x, y),
// 2 hypot(x, y, z) (x, y, z: the three input arrays).
void ref_jsm_geom(int32_t fn, const double* x, const double* y, const double* z, int64_t n, double* out)
{
   for (int64_t i = 0; i < n; ++i)
      out[i] = fn == 0 ? isc::jsm::cbrt(x[i]) : fn == 1 ? isc::jsm::hypot(x[i], y[i])
                                                        : isc::jsm::hypot(x[i], y[i], z[i]);
}

// ---- the golden cases' raw candidate cells (tests/parity's CellsCase / NpcsCase / LaminCase inputs) ----

static bool CaseCell(uint32_t seed, const char* const* names, const double* values, int32_t n, int32_t cx, int32_t cy,
                     isc::Params& p, isc::Cell& c)
{
   for (int32_t i = 0; i < n; ++i)
      if (!isc::SetParam(p, names[i], values[i]))
         return false;
   isc::NormalizeParams(p);
   c = isc::RawCandidate(seed, cx, cy, p);
   return c.present;
}

// The microtubules of the raw candidate cell (cx, cy) (BuildMicrotubulesForCell): points up to capPts, per-microtubule
// counts up to capMts, *nMts. Returns the point count, -1 when the chunk has no cell or a param is unknown.
int64_t ref_case_mts(uint32_t seed, const char* const* names, const double* values, int32_t n, int32_t cx, int32_t cy,
                     double* xyz, int64_t capPts, int32_t* lens, int32_t capMts, int32_t* nMts)
{
   isc::Params p;
   isc::Cell c;
   if (!CaseCell(seed, names, values, n, cx, cy, p, c))
      return -1;
   const std::vector<isc::Microtubule> mts = isc::BuildMicrotubulesForCell(seed, c, p, isc::BuildMtCellGeom(c, p));
   int64_t total = 0;
   for (size_t i = 0; i < mts.size(); ++i)
   {
      if ((int32_t)i < capMts)
         lens[i] = (int32_t)mts[i].pts.size();
      for (const isc::Pt3& q : mts[i].pts)
      {
         if (total < capPts)
         {
            xyz[3 * total] = q.x;
            xyz[3 * total + 1] = q.y;
            xyz[3 * total + 2] = q.z;
         }
         ++total;
      }
   }
   *nMts = (int32_t)mts.size();
   return total;
}

// The pores (which 1: rows c, n, e1, e2, cand, piece: 14 doubles; BuildNpcsForCell) or lamina points (which 2: rows
// c, n, cand, piece: 8 doubles; BuildLaminForCell) of the raw candidate cell (cx, cy), up to cap; sites / reachUm:
// NpcAttachmentPoints and NpcSiteReachUm of the params. Returns the count, -1 as ref_case_mts.
int32_t ref_case_envelope(uint32_t seed, const char* const* names, const double* values, int32_t n, int32_t cx,
                          int32_t cy, int32_t which, double* out, int32_t cap, double* sites, double* reachUm)
{
   isc::Params p;
   isc::Cell c;
   if (!CaseCell(seed, names, values, n, cx, cy, p, c))
      return -1;
   isc::NpcAttachmentPoints(p, sites);
   *reachUm = isc::NpcSiteReachUm(p);
   const std::vector<isc::Npc> pores = isc::BuildNpcsForCell(seed, cx, cy, c, p);
   if (which == 1)
   {
      for (size_t i = 0; i < pores.size() && (int32_t)i < cap; ++i)
      {
         const isc::Npc& q = pores[i];
         const double r[14] = {q.c.x,  q.c.y,  q.c.z,  q.n.x,  q.n.y,          q.n.z,          q.e1.x,
                               q.e1.y, q.e1.z, q.e2.x, q.e2.y, q.e2.z, (double)q.cand, (double)q.piece};
         std::memcpy(out + 14 * i, r, sizeof r);
      }
      return (int32_t)pores.size();
   }
   const std::vector<isc::LaminPoint> pts = isc::BuildLaminForCell(seed, cx, cy, c, p, pores);
   for (size_t i = 0; i < pts.size() && (int32_t)i < cap; ++i)
   {
      const isc::LaminPoint& q = pts[i];
      const double r[8] = {q.c.x, q.c.y, q.c.z, q.n.x, q.n.y, q.n.z, (double)q.cand, (double)q.piece};
      std::memcpy(out + 8 * i, r, sizeof r);
   }
   return (int32_t)pts.size();
}

// NpcDyesInBlock (which 1; rows as ref_case_envelope's, sites 64 x 3 nm) or LaminDyesInBlock (which 2) of the points
// rows[ids[0 .. nIds)]: rows pore, k, n, id, pos, site, rigid, tip (16 doubles), up to cap. Returns the count.
int32_t ref_envelope_dyes(int32_t which, uint32_t seed, int32_t cx, int32_t cy, const double* rows, int32_t nRows,
                          const int32_t* ids, int32_t nIds, const double* sites, double density, double ff,
                          const double* linkage4, double* out, int32_t cap)
{
   const isc::Linkage lk{linkage4[0], linkage4[1], linkage4[2], linkage4[3]};
   const std::vector<int32_t> idv(ids, ids + nIds);
   std::vector<isc::NpcDye> dyes;
   std::vector<isc::LinkageParts> parts;
   if (which == 1)
   {
      std::vector<isc::Npc> pores((size_t)nRows);
      for (int32_t i = 0; i < nRows; ++i)
      {
         const double* r = rows + 14 * i;
         isc::Npc& q = pores[(size_t)i];
         q.c = {r[0], r[1], r[2]};
         q.n = {r[3], r[4], r[5]};
         q.e1 = {r[6], r[7], r[8]};
         q.e2 = {r[9], r[10], r[11]};
         q.cand = (int32_t)r[12];
         q.piece = (int)r[13];
      }
      isc::NpcDyesInBlock(seed, cx, cy, pores, idv, sites, density, ff, dyes, lk, &parts);
   }
   else
   {
      std::vector<isc::LaminPoint> pts((size_t)nRows);
      for (int32_t i = 0; i < nRows; ++i)
      {
         const double* r = rows + 8 * i;
         isc::LaminPoint& q = pts[(size_t)i];
         q.c = {r[0], r[1], r[2]};
         q.n = {r[3], r[4], r[5]};
         q.cand = (int32_t)r[6];
         q.piece = (int)r[7];
      }
      isc::LaminDyesInBlock(seed, cx, cy, pts, idv, density, ff, dyes, lk, &parts);
   }
   for (size_t i = 0; i < dyes.size() && (int32_t)i < cap; ++i)
   {
      const isc::NpcDye& e = dyes[i];
      const isc::LinkageParts& P = parts[i];
      const double r[16] = {(double)e.pore, (double)e.k, (double)e.n, (double)e.id, e.pos.x,   e.pos.y,
                            e.pos.z,        P.site.x,    P.site.y,    P.site.z,     P.rigid.x, P.rigid.y,
                            P.rigid.z,      P.tip.x,     P.tip.y,     P.tip.z};
      std::memcpy(out + 16 * i, r, sizeof r);
   }
   return (int32_t)dyes.size();
}

// MtSiteGeometry (core/src/dyes.cpp:120) on a microtubule's points at arc S (um), azimuth theta and linker draws
// r1..r3: att, tip, dye (9 doubles).
void ref_mt_site_geometry(const double* xyz, int32_t nPts, double S, double theta, double r1, double r2, double r3,
                          double* out)
{
   std::vector<isc::Pt3> pts((size_t)nPts);
   for (int32_t i = 0; i < nPts; ++i)
      pts[(size_t)i] = isc::Pt3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
   const isc::MtFrames fr = isc::BuildMtFrames(pts);
   const isc::SiteGeom g = isc::MtSiteGeometry(pts, fr, isc::MtSegmentAt(fr, S), S, theta, r1, r2, r3);
   const double r[9] = {g.att.x, g.att.y, g.att.z, g.tip.x, g.tip.y, g.tip.z, g.dye.x, g.dye.y, g.dye.z};
   std::memcpy(out, r, sizeof r);
}

// MtProtofilamentTheta and MtProtofilamentOffsetNm (core/src/dyes.cpp:29-34).
void ref_mt_protofilament(double phase, int32_t k, double* theta, double* offNm)
{
   *theta = isc::MtProtofilamentTheta(phase, k);
   *offNm = isc::MtProtofilamentOffsetNm(k);
}

// World::GenerateDyes + Density3dInWindow's binning (core/src/world.cpp:808-838, 1519-1543) over caller-given
// microtubules, one thread: per cell (rows cx, cy, x, y, packRot), per microtubule (points xyz[3 ptStart[m] ..)),
// BuildMtFrames once, DyesInBlock for each 1 um block with the label (density, fraction, linkage), LocalToWorld, the
// dye store's 0.1 nm lattice, and the window's voxel counts (out, nx * ny * nz, added to). Returns the dyes counted.
int64_t ref_mt_density(uint32_t seed, const double* cells, int32_t nCells, const int32_t* mtStart,
                       const int64_t* ptStart, const double* xyz, double density, double ff, const double* linkage4,
                       const double* win, int32_t nx, int32_t ny, int32_t nz, float* out)
{
   const isc::Linkage lk{linkage4[0], linkage4[1], linkage4[2], linkage4[3]};
   const double sx = nx / (win[2] - win[0]), sy = ny / (win[3] - win[1]);
   const double sz = nz > 1 ? nz / (win[5] - win[4]) : 0.0;
   auto lattice = [](double um) {
      const int64_t g = (int64_t)std::floor(um * 10000.0 - 0.3819660112501051 + 0.5);
      return ((double)g + 0.3819660112501051) / 10000.0;
   };
   int64_t total = 0;
   std::vector<isc::Dye> dyes;
   for (int32_t c = 0; c < nCells; ++c)
   {
      const double* cr = cells + 5 * c;
      isc::Cell cell;
      cell.cx = (int32_t)cr[0];
      cell.cy = (int32_t)cr[1];
      cell.x = cr[2];
      cell.y = cr[3];
      cell.packRot = cr[4];
      for (int32_t m = mtStart[c]; m < mtStart[c + 1]; ++m)
      {
         std::vector<isc::Pt3> pts((size_t)(ptStart[m + 1] - ptStart[m]));
         for (size_t i = 0; i < pts.size(); ++i)
         {
            const double* q = xyz + 3 * (ptStart[m] + (int64_t)i);
            pts[i] = isc::Pt3{q[0], q[1], q[2]};
         }
         if (pts.size() < 2)
            continue;
         const isc::MtFrames fr = isc::BuildMtFrames(pts);
         const int nBlocks = (int)std::ceil(fr.Length() / isc::DYE_BLOCK_UM);
         for (int b = 0; b < nBlocks; ++b)
         {
            dyes.clear();
            isc::DyesInBlock(seed, cell.cx, cell.cy, m - mtStart[c], pts, fr, b, density, ff, dyes, lk);
            for (const isc::Dye& d : dyes)
            {
               double wx, wy;
               isc::LocalToWorld(cell, d.pos.x, d.pos.y, wx, wy);
               const double x = lattice(wx), y = lattice(wy), z = lattice(d.pos.z);
               if (!(z >= win[4] && z < win[5] && x >= win[0] && x < win[2] && y >= win[1] && y < win[3]))
                  continue;
               const int ix = std::min(nx - 1, (int)std::floor((x - win[0]) * sx));
               const int iy = std::min(ny - 1, (int)std::floor((y - win[1]) * sy));
               const int iz = nz > 1 ? std::min(nz - 1, (int)std::floor((z - win[4]) * sz)) : 0;
               out[((size_t)iz * ny + iy) * nx + ix] += 1;
               ++total;
            }
         }
      }
   }
   return total;
}

} // extern "C"
