// context.hip -- the iscg context: device selection, its stream, errors, host-argument staging.
#include "iscg_internal.h"

#include <cstdio>
#include <cstdlib>

namespace iscg {

void destroy_splat(iscg_context* h);
void destroy_psf(iscg_context* h);
void destroy_wf(iscg_context* h);
void destroy_roi(iscg_context* h);
void destroy_wf_accel(iscg_context* h);
void destroy_binned(iscg_context* h);
void destroy_wf_drift(iscg_context* h);

int32_t fail(iscg_context* h, const std::string& msg)
{
   if (h)
      h->err = msg;
   return -1;
}

bool ok(iscg_context* h, hipError_t e, const char* what)
{
   if (e == hipSuccess)
      return true;
   fail(h, std::string(what) + ": " + hipGetErrorString(e));
   return false;
}

int32_t begin(iscg_context* h)
{
   if (!h)
      return -1;
   h->err.clear();
   if (h->injectCount > 0)
   {
      --h->injectCount;
      return fail(h, h->injectMessage);
   }
   if (!ok(h, hipSetDevice(h->device), "hipSetDevice"))
      return -1;
   return 0;
}

int tuned_choice(iscg_context* h, const std::string& key, const std::vector<int>& cands, int dflt,
                 const std::function<bool(int)>& launch)
{
   if (h->tuneMode == 0 || cands.size() < 2)
      return dflt;
   const auto it = h->tuned.find(key);
   if (it != h->tuned.end())
      for (int c : cands)
         if (c == it->second)
            return c;
   hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
   if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
   {
      (void)hipGetLastError();
      return dflt;
   }
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
   {
      (void)hipGetLastError();
      if (e0)
         (void)hipEventDestroy(e0);
      return dflt;
   }
   int best = dflt;
   float bestMs = 0.0f;
   bool any = false, good = true;
   for (int c : cands)
   {
      if (!launch(c))
      {
         good = false;
         break;
      }
      if (hipEventRecord(e0, h->stream) != hipSuccess)
      {
         good = false;
         break;
      }
      for (int r = 0; r < kTuneReps && good; ++r)
         good = launch(c);
      float ms = 0.0f;
      if (!good || hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      {
         good = false;
         break;
      }
      if (!any || ms < bestMs)
      {
         best = c;
         bestMs = ms;
         any = true;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   if (!good)
   {
      (void)hipGetLastError();
      return dflt;
   }
   h->tuned[key] = best;
   return best;
}

bool is_device_ptr(const void* p)
{
   if (!p)
      return false;
   hipPointerAttribute_t a;
   const hipError_t e = hipPointerGetAttributes(&a, p);
   if (e != hipSuccess)
   {
      (void)hipGetLastError();
      return false;
   }
   return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

const void* stage_in(iscg_context* h, int slot, const void* p, size_t bytes)
{
   if (!p || bytes == 0 || is_device_ptr(p))
      return p;
   DevBuf& b = h->stage[slot];
   if (!ok(h, b.reserve(bytes), "staging allocation"))
      return nullptr;
   if (!ok(h, hipMemcpyAsync(b.p, p, bytes, hipMemcpyHostToDevice, h->stream), "host-to-device copy"))
      return nullptr;
   return b.p;
}

void* stage_out(iscg_context* h, int slot, void* p, size_t bytes)
{
   if (!p || bytes == 0 || is_device_ptr(p))
      return p;
   DevBuf& b = h->stage[slot];
   if (!ok(h, b.reserve(bytes), "staging allocation"))
      return nullptr;
   return b.p;
}

bool finish_out(iscg_context* h, void* user, const void* dev, size_t bytes)
{
   if (!user || bytes == 0 || user == dev)
      return true;
   if (!ok(h, hipMemcpyAsync(user, dev, bytes, hipMemcpyDeviceToHost, h->stream), "device-to-host copy"))
      return false;
   return ok(h, hipStreamSynchronize(h->stream), "stream synchronize");
}

static hipEvent_t take_event(iscg_context* h)
{
   if (!h->eventPool.empty())
   {
      hipEvent_t e = h->eventPool.back();
      h->eventPool.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   if (hipEventCreate(&e) != hipSuccess)
      return nullptr;
   return e;
}

ProfScope::ProfScope(iscg_context* hh, const char* nm) : h(hh), name(nm)
{
   if (h->profiling)
   {
      a = take_event(h);
      if (a && hipEventRecord(a, h->stream) != hipSuccess)
         a = nullptr;
   }
}

ProfScope::~ProfScope()
{
   if (!a)
      return;
   hipEvent_t b = take_event(h);
   if (b && hipEventRecord(b, h->stream) == hipSuccess)
      h->pending.push_back({name, a, b});
}

}  // namespace iscg

using namespace iscg;

namespace {
std::string g_createError;
}

extern "C" {

int32_t iscg_abi_version(void) { return ISCG_ABI_VERSION; }

int32_t iscg_create(int32_t device, iscg_context** out)
{
   if (!out)
      return -1;
   *out = nullptr;
   int n = 0;
   hipError_t e = hipGetDeviceCount(&n);
   if (e != hipSuccess || n <= 0)
   {
      g_createError = std::string("no HIP device: ") + hipGetErrorString(e);
      return -1;
   }
   if (device < 0 || device >= n)
   {
      g_createError = "device index out of range";
      return -1;
   }
   e = hipSetDevice(device);
   if (e != hipSuccess)
   {
      g_createError = std::string("hipSetDevice: ") + hipGetErrorString(e);
      return -1;
   }
   hipDeviceProp_t prop;
   e = hipGetDeviceProperties(&prop, device);
   if (e != hipSuccess)
   {
      g_createError = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
      return -1;
   }
   if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
   {
      g_createError = std::string("libiscg is built for gfx950; device is ") + prop.gcnArchName;
      return -1;
   }
   iscg_context* h = new iscg_context();
   h->device = device;
   e = hipStreamCreateWithFlags(&h->ownStream, hipStreamNonBlocking);
   h->stream = h->ownStream;
   if (e != hipSuccess)
   {
      g_createError = std::string("hipStreamCreate: ") + hipGetErrorString(e);
      delete h;
      return -1;
   }
   // the device address of the allocation-free noise entry point's scratch array, looked up once per device with
   // the device current; a failed look-up only disables that entry point (it reports it)
   std::string scratchErr;
   (void)noalloc_setup(device, scratchErr);
   g_createError.clear();
   *out = h;
   return 0;
}

void iscg_destroy(iscg_context* h)
{
   if (!h)
      return;
   (void)hipSetDevice(h->device);
   (void)hipStreamSynchronize(h->stream);
   (void)hipStreamSynchronize(h->ownStream);
   destroy_splat(h);
   destroy_psf(h);
   destroy_wf(h);
   destroy_wf_accel(h);
   destroy_roi(h);
   destroy_binned(h);
   destroy_wf_drift(h);
   for (auto& b : h->stage)
      b.release();
   for (auto& p : h->pending)
   {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
   }
   for (auto e : h->eventPool)
      (void)hipEventDestroy(e);
   if (h->noiseCount)
      (void)hipFree(h->noiseCount);
   (void)hipStreamDestroy(h->ownStream);
   delete h;
}

const char* iscg_last_error(const iscg_context* h) { return h ? h->err.c_str() : g_createError.c_str(); }

int32_t iscg_synchronize(iscg_context* h)
{
   if (!h)
      return -1;
   if (!ok(h, hipSetDevice(h->device), "hipSetDevice"))
      return -1;
   return ok(h, hipStreamSynchronize(h->stream), "stream synchronize") ? 0 : -1;
}

void* iscg_stream(iscg_context* h) { return h ? (void*)h->stream : nullptr; }

int32_t iscg_set_stream(iscg_context* h, void* stream)
{
   if (begin(h) != 0)
      return -1;
   hipStream_t s = (hipStream_t)stream;
   if (s == h->stream)
      return 0;
   // work already enqueued (uploads of kernels, plans, maps) completes before the new stream uses it
   if (!ok(h, hipStreamSynchronize(h->stream), "iscg_set_stream: synchronize"))
      return -1;
   h->stream = s;
   return 0;
}

int32_t iscg_profile_enable(iscg_context* h, int32_t on)
{
   if (!h)
      return -1;
   h->profiling = on != 0;
   return 0;
}

int32_t iscg_profile_read(iscg_context* h, char* names, int64_t* launches, double* totalMs, int32_t cap,
                          int32_t reset)
{
   if (!h)
      return -1;
   if (!ok(h, hipSetDevice(h->device), "hipSetDevice") || !ok(h, hipStreamSynchronize(h->stream), "synchronize"))
      return -1;
   for (auto& p : h->pending)
   {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
      {
         auto& t = h->profTotals[p.name];
         t.first += 1;
         t.second += ms;
      }
      h->eventPool.push_back(p.a);
      h->eventPool.push_back(p.b);
   }
   h->pending.clear();
   int32_t i = 0;
   for (const auto& kv : h->profTotals)
   {
      if (i < cap)
      {
         if (names)
         {
            std::strncpy(names + 64 * i, kv.first.c_str(), 63);
            names[64 * i + 63] = 0;
         }
         if (launches)
            launches[i] = kv.second.first;
         if (totalMs)
            totalMs[i] = kv.second.second;
      }
      ++i;
   }
   if (reset)
      h->profTotals.clear();
   return i;
}

int32_t iscg_autotune_set(iscg_context* h, int32_t mode)
{
   if (!h)
      return -1;
   if (mode != 0 && mode != 1)
      return fail(h, "iscg_autotune_set: mode is 0 (fixed defaults) or 1 (tune on first use)");
   h->tuneMode = mode;
   return 0;
}

int32_t iscg_autotune_dump(iscg_context* h, char* buf, int32_t cap)
{
   if (!h)
      return -1;
   std::string text;
   for (const auto& kv : h->tuned)
      text += kv.first + "=" + std::to_string(kv.second) + "\n";
   if (text.size() > 0x7FFFFFFFu)
      return fail(h, "iscg_autotune_dump: too many choices");
   if (buf && cap > 0)
   {
      const size_t n = text.size() < (size_t)cap - 1 ? text.size() : (size_t)cap - 1;
      std::memcpy(buf, text.data(), n);
      buf[n] = 0;
   }
   return (int32_t)text.size();
}

int32_t iscg_autotune_load(iscg_context* h, const char* text)
{
   if (!h)
      return -1;
   if (!text)
      return fail(h, "iscg_autotune_load: null text");
   std::map<std::string, int> add;
   const char* p = text;
   while (*p)
   {
      const char* e = std::strchr(p, '\n');
      const std::string line = e ? std::string(p, (size_t)(e - p)) : std::string(p);
      p = e ? e + 1 : p + line.size();
      if (line.empty())
         continue;
      const size_t eq = line.rfind('=');
      if (eq == std::string::npos || eq == 0 || eq + 1 == line.size())
         return fail(h, "iscg_autotune_load: a line is not key=value: " + line);
      char* end = nullptr;
      const long v = std::strtol(line.c_str() + eq + 1, &end, 10);
      if (*end != 0 || v < 0 || v > 1 << 20)
         return fail(h, "iscg_autotune_load: bad value in: " + line);
      add[line.substr(0, eq)] = (int)v;
   }
   for (const auto& kv : add)
      h->tuned[kv.first] = kv.second;
   return (int32_t)add.size();
}

int32_t iscg_inject_failure(iscg_context* h, int32_t count, const char* message)
{
   if (!h)
      return -1;
   h->injectCount = count < 0 ? 0 : count;
   h->injectMessage = message ? message : "injected failure";
   return 0;
}

}  // extern "C"
