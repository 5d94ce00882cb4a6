// wf_drift.hip -- the frames of a drifting sample on the mean field: RenderShiftedImages
// (Sim/WidefieldRender.cpp:2905-2935) for a batch of frames, the per-frame work of WidefieldDriftFrames::Render
// (:2964-2977).
//
// Frame f: its spectrum is, over its one or two focus images m (weights 1 - w and w), the float-coefficient sum of
// image m's kept channel spectra (persistent, then bleach j with a[f][j]) in the reference's order; times the phase
// ramp of the images' sub-cell pose minus the frame's drift (ApplyShiftRamp, Sim/DriftShift.cpp:24-48: FP64 phase,
// cosine at the Nyquist bins); the inverse transform cropped to the FOV (the row and column passes of
// wf_passes.h), then max(0, .) summed in FP64 over up x up cells into the camera image (ShiftedCamera,
// Sim/DriftShift.cpp:50-73; added to).
//
//   drift_ramp_kernel   per frame: the ramp's factors px[kx], py[ky] (FP64)          (W + NY) x 16 B per frame
//   drift_mac_kernel    per (frame, bin): the channel sum x ramp x 1 / (NX NY)          8 B per bin per term + 8 B out
//   inverse_xy          columns (FOV rows only) + rows C2R (FOV columns), raw values
//   drift_bin_kernel    per camera pixel: max(0, .) over up x up cells, FP64 sum, += cam
// All HBM-bound streaming; a batch of frames per launch (chunks of kDriftChunk frames bound the scratch).
#include "iscg_internal.h"
#include "wf_passes.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace iscg {

struct WfDriftState
{
   uint32_t NX = 0, NY = 0;
   DevBuf twX, twY, frames, coef, ramp, spec, x2, img;
};

void destroy_wf_drift(iscg_context* h)
{
   delete h->wfDrift;
   h->wfDrift = nullptr;
}

constexpr uint32_t kDriftChunk = 32;

// One term of a frame's spectrum sum: image m's channel c times a float coefficient.
struct DriftTerm
{
   int32_t image, channel;
   float coef;
   int32_t pad;
};

// ApplyShiftRamp's factors for frame f: px[kx] = e^{2 pi i kx fx / NX} (kx < W), py[ky] = e^{2 pi i ky' fy / NY}
// (ky' the signed index); the Nyquist bins take the cosine alone.
__global__ void drift_ramp_kernel(const double2* __restrict__ shift, uint32_t nFrames, uint32_t NX, uint32_t NY,
                                  double2* __restrict__ ramp)
{
   const uint32_t W = NX / 2 + 1, per = W + NY;
   const double kTwoPi = 6.283185307179586476925;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nFrames * per; i += gridDim.x * blockDim.x)
   {
      const uint32_t f = i / per, k = i % per;
      double a;
      bool nyq;
      if (k < W)
      {
         a = kTwoPi * (double)k * shift[f].x / (double)NX;
         nyq = k == NX / 2;
      }
      else
      {
         const uint32_t ky = k - W;
         const long kys = ky <= NY / 2 ? (long)ky : (long)ky - (long)NY;
         a = kTwoPi * (double)kys * shift[f].y / (double)NY;
         nyq = ky == NY / 2;
      }
      double s, c;
      sincos(a, &s, &c);
      ramp[i] = make_double2(c, nyq ? 0.0 : s);
   }
}

// out[f][ky][kx] = (sum_t coef_t spec[image_t][channel_t][ky][kx]) x px[kx] py[ky] x norm (blockIdx.y = frame)
__global__ __launch_bounds__(256) void drift_mac_kernel(const float2* __restrict__ spec, size_t imageStride,
                                                        size_t channelStride, const DriftTerm* __restrict__ terms,
                                                        const uint32_t* __restrict__ termStart,
                                                        const double2* __restrict__ ramp, uint32_t NX, uint32_t NY,
                                                        uint32_t S, double norm, float2* __restrict__ out)
{
   const uint32_t f = blockIdx.y;
   const uint32_t W = NX / 2 + 1, bins = NY * W;
   const uint32_t t0 = termStart[f], t1 = termStart[f + 1];
   const double2* px = ramp + (size_t)f * (W + NY);
   const double2* py = px + W;
   float2* o = out + (size_t)f * NY * S;
   for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x)
   {
      const uint32_t ky = b / W, kx = b % W;
      const size_t at = (size_t)ky * W + kx;
      float2 acc = make_float2(0.0f, 0.0f);
      for (uint32_t t = t0; t < t1; ++t)
      {
         const DriftTerm d = terms[t];
         const float2 v = spec[(size_t)d.image * imageStride + (size_t)d.channel * channelStride + at];
         acc.x += d.coef * v.x;  // S += c x spec, as the reference's complex<float> sum
         acc.y += d.coef * v.y;
      }
      // complex<double>(v) x (px py), then to float; the transform's 1 / (NX NY) folded in
      const double2 a = px[kx], c = py[ky];
      const double rr = a.x * c.x - a.y * c.y, ri = a.x * c.y + a.y * c.x;
      const double vr = (double)acc.x * rr - (double)acc.y * ri, vi = (double)acc.x * ri + (double)acc.y * rr;
      o[(size_t)ky * S + kx] = make_float2((float)(vr * norm), (float)(vi * norm));
   }
   for (uint32_t b = W + blockIdx.x * blockDim.x + threadIdx.x; b < S; b += gridDim.x * blockDim.x)
      for (uint32_t ky = 0; ky < NY; ++ky)
         o[(size_t)ky * S + b] = make_float2(0.0f, 0.0f);
}

// cam[f][Y][X] += (float)(sum over the up x up cells of max(0, img) in FP64 x scale)
__global__ void drift_bin_kernel(const float* __restrict__ img, uint32_t nFrames, uint32_t cw, uint32_t ch,
                                 uint32_t up, float scale, float* __restrict__ cam)
{
   const uint32_t Wc = cw / up, Hc = ch / up, per = Wc * Hc;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nFrames * per; i += gridDim.x * blockDim.x)
   {
      const uint32_t f = i / per, r = i % per, Y = r / Wc, X = r % Wc;
      const float* src = img + (size_t)f * cw * ch;
      double acc = 0.0;
      for (uint32_t sy = 0; sy < up; ++sy)
         for (uint32_t sx = 0; sx < up; ++sx)
         {
            const float v = src[(size_t)(Y * up + sy) * cw + X * up + sx];
            acc += v > 0.0f ? v : 0.0f;
         }
      cam[i] += (float)(acc * scale);
   }
}

}  // namespace iscg

using namespace iscg;

extern "C" int32_t iscg_wf_shifted_frames(iscg_context* h, const iscg_wf_shift_desc* d, const float* spectra,
                                          const iscg_wf_shift_frame* frames, const double* a, uint32_t nFrames,
                                          float* cam)
{
   if (begin(h))
      return -1;
   if (!d || (nFrames && (!spectra || !frames || !cam)))
      return fail(h, "iscg_wf_shifted_frames: null argument");
   if (!pow2(d->NX, 2048) || !pow2(d->NY, 2048))
      return fail(h, "iscg_wf_shifted_frames: FFT grid must be powers of two in [16, 2048]");
   const uint32_t up = d->upscale < 1 ? 1 : (uint32_t)d->upscale;
   if (d->cw == 0 || d->ch == 0 || d->fovX0 + d->cw > d->NX || d->fovY0 + d->ch > d->NY || d->cw % up || d->ch % up)
      return fail(h, "iscg_wf_shifted_frames: the FOV must lie in the grid and bin by the upscale");
   const int nCh = d->nChannels, hasP = d->hasPersistent ? 1 : 0;
   if (d->nImages < 1 || nCh < hasP || nCh < 1)
      return fail(h, "iscg_wf_shifted_frames: no images or channels");
   const int nBleach = nCh - hasP;
   for (uint32_t f = 0; f < nFrames; ++f)
   {
      const iscg_wf_shift_frame& fr = frames[f];
      if (fr.i0 < 0 || fr.i0 >= d->nImages || fr.i1 >= d->nImages)
         return fail(h, "iscg_wf_shifted_frames: frame " + std::to_string(f) + " names an image outside the set");
   }
   if (nFrames == 0)
      return 0;
   if (!h->wfDrift)
      h->wfDrift = new WfDriftState();
   WfDriftState& s = *h->wfDrift;
   if (s.NX != d->NX || s.NY != d->NY)
   {
      std::vector<float2> t;
      auto tw = [&](DevBuf& b, uint32_t n) {
         t.resize(n);
         for (uint32_t k = 0; k < n; ++k)
         {
            const double ang = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
            t[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
         }
         return ok(h, b.reserve(n * sizeof(float2)), "twiddle allocation") &&
                ok(h, hipMemcpyAsync(b.p, t.data(), n * sizeof(float2), hipMemcpyHostToDevice, h->stream),
                   "twiddle copy") &&
                ok(h, hipStreamSynchronize(h->stream), "twiddle copy");
      };
      if (!tw(s.twX, d->NX) || !tw(s.twY, d->NY))
         return -1;
      s.NX = d->NX;
      s.NY = d->NY;
   }
   const uint32_t NX = d->NX, NY = d->NY, W = NX / 2 + 1, S = spec_stride(NX);
   const size_t chStride = (size_t)NY * W, imStride = chStride * nCh;
   const float2* dSpec = (const float2*)stage_in(h, 0, spectra, imStride * d->nImages * sizeof(float2));
   const size_t camPer = (size_t)(d->cw / up) * (d->ch / up);
   // the camera images are added to: a host array is staged in and copied back
   float* dCam = (float*)stage_in(h, 1, cam, camPer * nFrames * sizeof(float));
   if (!dSpec || !dCam)
      return -1;
   // per frame: its terms in the reference's order (i0's persistent, i0's bleach, then i1's) and its shift
   std::vector<DriftTerm> terms;
   std::vector<uint32_t> start{0};
   std::vector<double2> shift(nFrames);
   for (uint32_t f = 0; f < nFrames; ++f)
   {
      const iscg_wf_shift_frame& fr = frames[f];
      const bool two = fr.i1 >= 0 && fr.w != 0.0;
      auto addImage = [&](int m, double wi) {
         const float cwt = (float)wi;
         if (hasP && cwt != 0.0f)
            terms.push_back({m, 0, cwt, 0});
         for (int j = 0; j < nBleach; ++j)
         {
            const double aj = a ? a[(size_t)f * nBleach + j] : 0.0;
            const float c = (float)aj * cwt;
            if (aj != 0.0 && c != 0.0f)
               terms.push_back({m, hasP + j, c, 0});
         }
      };
      addImage(fr.i0, fr.i1 >= 0 ? 1.0 - fr.w : 1.0);
      if (two)
         addImage(fr.i1, fr.w);
      start.push_back((uint32_t)terms.size());
      shift[f] = make_double2(d->fracX - fr.dxCells, d->fracY - fr.dyCells);
   }
   if (terms.empty())
      terms.push_back({0, 0, 0.0f, 0});
   if (!ok(h, s.coef.reserve(terms.size() * sizeof(DriftTerm) + (start.size() + 64) * sizeof(uint32_t)), "alloc") ||
       !ok(h, s.frames.reserve(shift.size() * sizeof(double2)), "alloc") ||
       !ok(h, s.ramp.reserve((size_t)kDriftChunk * (W + NY) * sizeof(double2)), "alloc") ||
       !ok(h, s.spec.reserve((size_t)kDriftChunk * NY * S * sizeof(float2)), "alloc") ||
       !ok(h, s.x2.reserve((size_t)kDriftChunk * d->ch * S * sizeof(float2)), "alloc") ||
       !ok(h, s.img.reserve((size_t)kDriftChunk * d->ch * d->cw * sizeof(float)), "alloc"))
      return -1;
   DriftTerm* dTerms = s.coef.as<DriftTerm>();
   uint32_t* dStart = (uint32_t*)(dTerms + terms.size());
   if (!ok(h, hipMemcpyAsync(dTerms, terms.data(), terms.size() * sizeof(DriftTerm), hipMemcpyHostToDevice,
                             h->stream), "copy") ||
       !ok(h, hipMemcpyAsync(dStart, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream),
           "copy") ||
       !ok(h, hipMemcpyAsync(s.frames.p, shift.data(), shift.size() * sizeof(double2), hipMemcpyHostToDevice,
                             h->stream), "copy"))
      return -1;
   const XyGrid grid{NX, NY, S, s.twX.as<const float2>(), s.twY.as<const float2>()};
   for (uint32_t f0 = 0; f0 < nFrames; f0 += kDriftChunk)
   {
      const uint32_t nf = std::min(kDriftChunk, nFrames - f0);
      {
         ProfScope ps(h, "wf_drift_mac");
         hipLaunchKernelGGL(drift_ramp_kernel, dim3(grid_for((size_t)nf * (W + NY), 256, 1)), dim3(256), 0,
                            h->stream, s.frames.as<const double2>() + f0, nf, NX, NY, s.ramp.as<double2>());
         const unsigned gx = (unsigned)std::min<uint32_t>((NY * W + 255) / 256, 1024);
         hipLaunchKernelGGL(drift_mac_kernel, dim3(gx, nf), dim3(256), 0, h->stream, dSpec, imStride, chStride,
                            dTerms, dStart + f0, s.ramp.as<const double2>(), NX, NY, S, 1.0 / ((double)NX * NY),
                            s.spec.as<float2>());
      }
      C2rEpilogue e{};
      e.raw = 1;
      e.img = s.img.as<float>();
      e.ch = (int)d->ch;
      e.cw = (int)d->cw;
      e.x0 = (int)d->fovX0;
      if (!inverse_xy(h, grid, s.spec.as<const float2>(), (int)nf, s.x2.as<float2>(), (int)d->fovY0, e,
                      "wf_drift_c2r"))
         return fail(h, "iscg_wf_shifted_frames: unsupported FFT size");
      {
         ProfScope ps(h, "wf_drift_bin");
         hipLaunchKernelGGL(drift_bin_kernel, dim3(grid_for(camPer * nf, 256, 4)), dim3(256), 0, h->stream,
                            s.img.as<const float>(), nf, d->cw, d->ch, up, 1.0f, dCam + camPer * f0);
      }
      if (!ok(h, hipGetLastError(), "iscg_wf_shifted_frames launch"))
         return -1;
   }
   if (!finish_out(h, cam, dCam, camPer * nFrames * sizeof(float)))
      return -1;
   return 0;
}
