// WidefieldIscgHost.h -- a sim::WidefieldAccelerator (Sim/WidefieldRender.h:342-361) on libiscg: the mean-field
// scenes' per-focus work (WidefieldScene in GPU mode) on MI355X.
//
// The accelerator is the reference's abstract seam, so this header is all a host adds: it subclasses it against the
// unmodified Sim/WidefieldRender.h and flattens each WidefieldGpuJob into iscg_wf_job (include/iscg.h). Install it
// where the reference installs its Direct3D 11 host: WidefieldScene::SetAccelerator(&host) with the scene in GPU mode
// (SetGpuMode(true): power-of-two grids), or FluorescenceMovie::Begin(spec, true, err, &host) (Sim/ScopeMovie.h:433).
// A failed Images leaves err set and the scene renders on the CPU from then on, as with the reference's hosts.
#pragma once

#include "WidefieldRender.h"

#include "../../include/iscg.h"

#include <cstdint>
#include <string>
#include <vector>

namespace sim {

class WidefieldIscgHost : public WidefieldAccelerator
{
public:
   // h: a context the host uses alone (not re-entrant); the caller keeps it alive and destroys it.
   explicit WidefieldIscgHost(iscg_context* h) : h_(h) {}

   bool HasPlane(unsigned long long geometry, unsigned long long key) const override
   {
      return iscg_wf_has_plane(h_, geometry, key) == 1;
   }

   bool Fits(unsigned long long geometry, unsigned NX, unsigned NY, size_t planes) override
   {
      return iscg_wf_fits(h_, geometry, NX, NY, planes) == 1;
   }

   bool Images(const WidefieldGpuJob& job, std::vector<std::vector<float>>& out, std::string& err) override
   {
      kp_.clear();
      ks_.clear();
      for (const auto& k : job.kernels)
      {
         kp_.push_back(k.p);
         ks_.push_back(reinterpret_cast<const float*>(k.spec->data()));  // cfloat: (re, im) floats
      }
      keys_.clear();
      starts_.assign(1, 0u);
      cells_.clear();
      values_.clear();
      for (const auto& p : job.planes)
      {
         keys_.push_back(p.key);
         cells_.insert(cells_.end(), p.cells.begin(), p.cells.end());
         values_.insert(values_.end(), p.values.begin(), p.values.end());
         starts_.push_back(static_cast<uint32_t>(cells_.size()));
      }
      chan_.assign(1, 0u);
      deps_.clear();
      for (const auto& ch : job.channels)
      {
         for (const auto& d : ch)
            deps_.push_back(iscg_wf_dep{d.key, d.p0, d.p1, d.w0, d.w1, d.two ? 1 : 0, 0});
         chan_.push_back(static_cast<uint32_t>(deps_.size()));
      }
      iscg_wf_job j;
      j.NX = job.NX;
      j.NY = job.NY;
      j.nx = job.nx;
      j.ny = job.ny;
      j.fovX0 = job.fovX0;
      j.fovY0 = job.fovY0;
      j.cw = job.cw;
      j.ch = job.ch;
      j.fracX = job.fracX;
      j.fracY = job.fracY;
      j.geometry = job.geometry;
      j.nKernels = static_cast<int32_t>(kp_.size());
      j.nPlanes = static_cast<int32_t>(keys_.size());
      j.nChannels = static_cast<int32_t>(job.channels.size());
      j.hasPersistent = job.hasPersistent ? 1 : 0;
      j.kernelP = kp_.data();
      j.kernelSpec = ks_.data();
      j.planeKey = keys_.data();
      j.planeStart = starts_.data();
      j.cells = cells_.data();
      j.values = values_.data();
      j.channelStart = chan_.data();
      j.deps = deps_.data();
      const size_t n = static_cast<size_t>(job.cw) * job.ch;
      flat_.resize(n * job.channels.size());
      if (iscg_wf_images(h_, &j, flat_.data()) != 0)
      {
         err = iscg_last_error(h_);
         return false;
      }
      out.assign(job.channels.size(), std::vector<float>());
      for (size_t c = 0; c < out.size(); ++c)
         out[c].assign(flat_.begin() + c * n, flat_.begin() + (c + 1) * n);
      return true;
   }

private:
   iscg_context* h_;
   // the flattened job, kept between calls so a refocus allocates nothing new on the host
   std::vector<int32_t> kp_;
   std::vector<const float*> ks_;
   std::vector<uint64_t> keys_;
   std::vector<uint32_t> starts_, cells_, chan_;
   std::vector<float> values_, flat_;
   std::vector<iscg_wf_dep> deps_;
};

}  // namespace sim
