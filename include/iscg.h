/*
 * iscg.h -- C ABI of the MI355X image-formation library (libiscg.so, gfx950).
 *
 * The entry points sit under the reference renderer's existing C++ seams (Sim/ = adapter/inSiliScope/Simulation/):
 * each one names the reference interface it replaces. Plain pointers and sizes only: no C++ or torch types.
 *
 * Conventions (following the core's C ABI, core/include/insiliscope/insiliscope.h:1-60):
 *   - every call returns int32_t: 0 on success, < 0 on failure; iscg_last_error() then says why;
 *   - a context is one HIP device and one HIP stream; calls on one context are serialised by the caller
 *     (not re-entrant), as GpuSimulator is one instance per thread (Sim/GpuSimD3D11.h:28-30);
 *   - array arguments may be host memory or device memory of the context's device: device pointers are used in
 *     place (the fast path: inputs resident in HBM), host pointers are staged through device buffers the context
 *     owns. Calls return after their results are written (host outputs) or enqueued on the context's stream
 *     (device outputs: iscg_synchronize, or events on iscg_stream(), orders later work);
 *   - the arithmetic matches the reference CPU path: the counter RNG and camera noise in FP64 (bit-exact ADU),
 *     the splat in FP64 accumulation (bit-exact photon images), the PSF and mean field in FP32 / split precision
 *     checked against FP64 to <= 1e-5 relative L2.
 */
#ifndef ISCG_H
#define ISCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISCG_ABI_VERSION 1

typedef struct iscg_context iscg_context;

/* ---- context ------------------------------------------------------------------------------------------------- */

int32_t iscg_abi_version(void);
/* A context on HIP device `device` with its own default-priority stream. */
int32_t iscg_create(int32_t device, iscg_context** out);
void iscg_destroy(iscg_context* h);
/* The last failure's message (empty after a success); valid until the next call on h. NULL h: creation errors. */
const char* iscg_last_error(const iscg_context* h);
/* Waits for everything enqueued on the context's stream. */
int32_t iscg_synchronize(iscg_context* h);
/* The context's hipStream_t (as void*), for callers that time or order work with HIP events. */
void* iscg_stream(iscg_context* h);
/* Enqueues the context's later work on `stream`, a hipStream_t of the context's device (NULL: the null stream), so
 * that a caller's own stream orders it with the caller's kernels; the current stream is synchronized first. The
 * stream stays the caller's: iscg_destroy does not destroy it. iscg_stream() before any call returns the context's
 * own stream, which may be passed back. No reference counterpart (the reference's GPU path owns its device queue). */
int32_t iscg_set_stream(iscg_context* h, void* stream);
/* Test hook: the next `count` entry-point calls on h fail with `message` before doing any work, as a device error
 * would (a caller's CPU fallback, Sim/ScopeMovie.cpp:7634-7645, is exercised without faulting a GPU). */
int32_t iscg_inject_failure(iscg_context* h, int32_t count, const char* message);

/* Launch-parameter autotuning. No reference counterpart: the reference's GPU paths use fixed tilings. Some launch
 * parameters have several values that give bit-identical results (the persistent splat grid's workgroups per CU, the
 * workgroups that share one PSF request's planes). With mode 1 (the default) the first call with a new shape key
 * (sizes, batch size rounded up to a power of two) times every candidate on the context's stream, one warm-up and 3
 * timed launches each, and keeps the fastest for that key; mode 0 uses the fixed defaults. No tuning happens while
 * the stream is being captured into a graph (defaults, nothing kept), so tune with an eager call first. */
int32_t iscg_autotune_set(iscg_context* h, int32_t mode);
/* The kept choices as text, one "key=value" line each, into buf (cap bytes, NUL-terminated, may be NULL); returns
 * the length of the whole text. iscg_autotune_load adds the choices of such a text (a cache kept across processes,
 * as the reference keeps PSF kernels on disk) and returns how many; a kept value that is not a candidate of this
 * build is tuned again on use. */
int32_t iscg_autotune_dump(iscg_context* h, char* buf, int32_t cap);
int32_t iscg_autotune_load(iscg_context* h, const char* text);

/* The movie file (replaces TiffWriter, cli/tiff_writer.h:24-199; host code, no context): little-endian baseline
 * TIFF, one uncompressed grey page per frame, the reference's layout byte for byte. open: BigTIFF when bigTiff != 0
 * or expectedBytes + 64 MiB >= 2^32 - 1 (iscg_tiff_bytes gives the expected size); the writer exists even when
 * open fails (read iscg_tiff_error, then destroy). page_u16 / page_f32: a 16-bit unsigned or 32-bit float page of
 * w x h samples with an optional description (NULL or ""); page_rows_u16: a page whose first `rows` rows are given,
 * the rest written later by rows_u16 (rows y0 .. y0 + rows - 1 of a 0-based page). frames_u16: nFrames pages from
 * host or device memory, the description on the first. A classic file that would pass 4 GB is refused. close
 * flushes and closes (0, or -1 with the error kept); destroy frees. Every call returns 0 or -1. */
typedef struct iscg_tiff iscg_tiff;
uint64_t iscg_tiff_bytes(uint64_t w, uint64_t h, uint64_t pages, uint64_t bytesPerSample, uint64_t descBytes);
int32_t iscg_tiff_open(const char* path, uint64_t expectedBytes, int32_t bigTiff, iscg_tiff** out);
int32_t iscg_tiff_page_u16(iscg_tiff* t, const uint16_t* px, uint32_t w, uint32_t h, const char* desc);
int32_t iscg_tiff_page_f32(iscg_tiff* t, const float* px, uint32_t w, uint32_t h, const char* desc);
int32_t iscg_tiff_page_rows_u16(iscg_tiff* t, const uint16_t* px, uint32_t rows, uint32_t w, uint32_t h,
                                const char* desc);
int32_t iscg_tiff_rows_u16(iscg_tiff* t, uint64_t page, uint32_t y0, uint32_t rows, const uint16_t* px);
int32_t iscg_tiff_frames_u16(iscg_tiff* t, const uint16_t* frames, uint32_t nFrames, uint32_t w, uint32_t h,
                             const char* firstDesc);
int32_t iscg_tiff_info(const iscg_tiff* t, uint64_t* pages, uint64_t* bytes, int32_t* big);
const char* iscg_tiff_error(const iscg_tiff* t);
int32_t iscg_tiff_close(iscg_tiff* t);
void iscg_tiff_destroy(iscg_tiff* t);

/* Per-kernel timing on the context's stream (the ISC_PROFILE analogue, Sim/Timing.h:1-60): with profiling on, each
 * kernel launch is bracketed by HIP events. iscg_profile_read waits for the stream and reports, per kernel name,
 * launches and total milliseconds (names: cap x 64 chars); returns the number of kernels. reset != 0 clears. */
int32_t iscg_profile_enable(iscg_context* h, int32_t on);
int32_t iscg_profile_read(iscg_context* h, char* names, int64_t* launches, double* totalMs, int32_t cap,
                          int32_t reset);

/* ---- the core's counter hash (core/src/rng.h:18-52) --------------------------------------------------------------- */

/* Pcg4d of n addresses (a[i], b[i], c[i], d[i]): words[4 i .. 4 i + 3] = the four output words. */
int32_t iscg_pcg4d(iscg_context* h, const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* d,
                   uint32_t n, uint32_t* words);
/* HashUnit (core/src/rng.h:40-44) of n addresses: out[i] = ((Pcg4d(seed, cx[i], cy[i], k[i]).a >> 9) + 0.5) 2^-23. */
int32_t iscg_hash_units(iscg_context* h, uint32_t seed, const int32_t* cx, const int32_t* cy, const uint32_t* k,
                        uint32_t n, double* out);
/* The first n draws of HashStream(seed, cx, cy, base) (core/src/rng.h:46-52): HashUnit(seed, cx, cy, base * 4096 + i),
 * uint32 wrap. */
int32_t iscg_hash_stream(iscg_context* h, uint32_t seed, int32_t cx, int32_t cy, uint32_t base, uint32_t n,
                         double* out);

/* ---- camera noise ---------------------------------------------------------------------------------------------- */

/* sim::CameraNoiseParams (Sim/SMLMNoise.h:35-50). */
typedef struct
{
   double quantumEfficiency;
   double darkCurrentElectrons;
   double gainPhotonsPerAdu;
   double readNoiseElectrons;
   int32_t emccd;
   int32_t bitDepth;
   double emGain;
   double cicElectrons;
} iscg_camera;

/* The evaluation path of the camera noise (iscg_noise, the splat and mean-field noise epilogues). No reference
 * counterpart: both paths give the ADU of sim::ApplyNoiseChain. exact = 1 evaluates every pixel with the FP64 chain;
 * exact = 0 (the default) evaluates the Box-Muller normal and the Poisson threshold exp(-l) in FP32, takes each
 * decision (a count, the final rounding) only where it holds over twice their error bound, and evaluates the other
 * pixels with the FP64 chain. The bounds are verified on the device once per process, over every input the FP32
 * functions can receive; a device that fails the check uses the FP64 chain. */
int32_t iscg_noise_set_path(iscg_context* h, int32_t exact);

/* out[0] = 1 when h uses the FP32 path, else 0; out[1] = the verified bound on |g_fp32 - g| of one normal;
 * out[2] = the verified relative bound of the FP32 threshold. Runs the verification if it has not run. */
int32_t iscg_noise_bounds(iscg_context* h, double* out);

/* Replaces sim::ApplyNoiseChain (Sim/SMLMNoise.h:136, Sim/SMLMNoise.cpp:142-216), batched over frames.
 * photons: nFrames x height x width floats; frame k's stream is (seed, frames[k], pixel + pixelBase).
 * offset / gain / readNoise: width*height maps or NULL (NULL: the scalar of cam; a NULL offset map is offset 0,
 * as an empty PixelOffsetMap). out: nFrames x height x width uint16 ADU. */
int32_t iscg_noise(iscg_context* h, const float* photons, uint32_t width, uint32_t height, const iscg_camera* cam,
                   const float* offset, const float* gain, const float* readNoise, uint32_t seed,
                   const uint32_t* frames, uint32_t nFrames, uint32_t pixelBase, uint16_t* out);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): iscg_noise on device arrays only, allocating nothing. Its
 * deferred-pixel queue lives in a fixed device array of 256 B + 8 B x ISCG_NOISE_NOALLOC_MAX_PIXELS (32 MiB per
 * device, held until the process ends); a call takes at most ISCG_NOISE_NOALLOC_MAX_PIXELS pixels (width x height x
 * nFrames) and fails with an error above that or when any array is host memory. Same ADU as iscg_noise. Calls on one
 * GPU share the array: calls that overlap on one GPU (other streams, other contexts, other threads) race. */
#define ISCG_NOISE_NOALLOC_MAX_PIXELS (1u << 22)
int32_t iscg_noise_noalloc(iscg_context* h, const float* photons, uint32_t width, uint32_t height,
                           const iscg_camera* cam, const float* offset, const float* gain, const float* readNoise,
                           uint32_t seed, const uint32_t* frames, uint32_t nFrames, uint32_t pixelBase, uint16_t* out);

/* The camera's static maps and noise seeds for a random seed (host code; no context): NoiseSetup
 * (Sim/ScopeMovie.cpp:973-996) with PixelOffsetMap / PixelGainMap / PixelReadNoiseMap::Generate and GaussianRng
 * (Sim/SMLMNoise.cpp:17-140): one std::mt19937_64(seed), offset then gain then read-noise maps, each over the whole
 * sensor (sensorHeight rows, or height when sensorHeight is 0), one Box-Muller normal per pixel; the gain floored
 * at 4 % of nominal, the read noise at 0. With a window (sensorHeight != 0) the maps hold rows rowY0 .. rowY0 +
 * height - 1 and *pixelBase = rowY0 * width. *noiseSeed = seed ^ 0x9E3779B9. Each map: width x height floats. */
int32_t iscg_noise_setup(int64_t seed, uint32_t width, uint32_t height, uint32_t sensorHeight, uint32_t rowY0,
                         double offsetAdu, double offsetStdAdu, double gainPhotonsPerAdu, double gainStdFraction,
                         double readNoiseElectrons, double readStdFraction, float* offset, float* gain,
                         float* readNoise, uint32_t* noiseSeed, uint32_t* pixelBase);

/* ---- blink splat (+ noise) -------------------------------------------------------------------------------------- */

/* One kernel stack the splat reads (a PsfKernelCache, Sim/PsfGeneratorBridge.h:243-293): its sum-1 planes; the
 * library builds the block sums (PsfKernelPlanes::Sums, Sim/PsfGeneratorBridge.cpp:1069-1088) on the device. */
typedef struct
{
   const float* planes;   /* nz x size x size, row-major, each summing to 1 */
   const int16_t* spanLo; /* halo cut (PsfHaloSpans::lo/hi, nz x (2 camRad + 1)), or NULL: the whole square */
   const int16_t* spanHi;
   double zStepNm;
   uint64_t serial;       /* PsfKernelCache::SplatSerial(); equal non-zero serials skip the re-upload */
   int32_t nz;
   int32_t size;          /* sizeOversampled = 2 halfWidthOversampled + 1 */
   int32_t oversampling;
   int32_t interpMode;    /* PsfInterpMode: 0 nearest, 1 linear, 2 cubic (Catmull-Rom on block sums), 3 Fft (one
                           * Fourier shift of the raw plane per emitter, PlanSplat :1443-1456; render_frames only,
                           * size <= 255) */
} iscg_kernel_view;

/* One emitter of a frame as RenderPhotonImage places it (sim::FrameEmitter, Sim/SMLMSimulation.h:158): position in
 * camera pixels, photons for the frame, its z plane (NearestZIndex) and kernel index. The splat setup is computed
 * on the device in FP64, so frames are bit-exact with the CPU. 32 bytes. */
typedef struct
{
   double xPx, yPx, photons;
   int32_t zIndex, kernel;
} iscg_frame_emitter;

/* sim::GpuSplatEmitter (Sim/GpuSimD3D11.h:50-59), the 64-byte wire format CollectGpuEmitters makes
 * (Sim/SMLMSimulation.cpp:219): setup done on the host, float weights. */
typedef struct
{
   int32_t x0, y0, bx, by;
   int32_t plane; /* z plane | kernel index << 16 */
   int32_t nTaps;
   float photons;
   int32_t cutRadius2;
   float wx[4];
   float wy[4];
} iscg_splat_emitter;

/* Replaces WithHaloCut / HaloPlaneSpans (Sim/PsfGeneratorBridge.cpp:1137-1243): the halo cut of a kernel stack
 * (nz sum-1 planes of size x size, host or device) at threshold (a pixel's share of an emitter's photons; Cubic and
 * Fft reads keep a 1.25^2 margin). Per plane and dy in [-camRad, camRad], camRad = ((size - 1) / 2) / os: the
 * first and last kept dx into lo, hi (nz x (2 camRad + 1) int16, host or device; lo > hi: none) and the largest
 * kept dx^2 + dy^2 into radius2 (nz int32; -1: none); keptFraction (may be NULL) the kept share of the squares.
 * Returns camRad, 0 when the reference makes no cut (threshold <= 0 or camRad < 1: outputs untouched), -1 on
 * error. Synchronizes the context's stream. */
int32_t iscg_halo_spans(iscg_context* h, const float* planes, int32_t nz, int32_t size, int32_t oversampling,
                        int32_t interpMode, double threshold, int16_t* lo, int16_t* hi, int32_t* radius2,
                        double* keptFraction);
/* Replaces SplatGpuHost::SetKernels / GpuSimulator::SetKernels (Sim/SplatGpu.h:50, Sim/GpuSimD3D11.h:75):
 * up to 4 kernels, resident until replaced. */
int32_t iscg_splat_set_kernels(iscg_context* h, const iscg_kernel_view* kernels, int32_t n);

/* Replaces SplatGpuHost::SetStatic / GpuSimulator::SetStatic (Sim/SplatGpu.h:53, Sim/GpuSimD3D11.h:83): frame size,
 * the window's first sensor pixel (noise stream), camera maps (width*height or NULL) and the background: the map
 * (photons/pixel/frame, NULL: flatBackground everywhere). key != 0 equal to the last call's: kept as they are. */
int32_t iscg_splat_set_static(iscg_context* h, uint32_t width, uint32_t height, uint32_t pixelBase,
                              const float* offset, const float* gain, const float* readNoise, const float* background,
                              double flatBackground, uint64_t key);

/* A blink event as the splat's caller holds it (sim::BlinkEvent, Sim/SMLMSimulation.h:82-110: the fields the
 * image uses): position in um, height z in nm, [tStart, tEnd) in frames, brightness and excitation factors. 56 bytes. */
typedef struct
{
   double xUm, yUm, zNm, tStart, tEnd, brightness, excite;
} iscg_blink_event;

/* Replaces BucketEventsByFrame (Sim/SMLMSimulation.cpp:203-217) on the device: for each frame f < nFrames the
 * indices of the events overlapping it (tEnd > 0, tStart < nFrames, f in [floor(tStart), floor(tEnd)] clipped to
 * the movie), in event order: frame f's list is bucketIdx[bucketStart[f] .. bucketStart[f+1]) (nFrames + 1 offsets).
 * total receives the number of entries; more than idxCap fails (with total set, so the caller can size the output).
 * Arrays may be host or device memory; the call synchronizes the context's stream. */
int32_t iscg_bucket_events(iscg_context* h, const iscg_blink_event* events, uint32_t nEvents, uint32_t nFrames,
                           uint32_t* bucketStart, uint32_t* bucketIdx, uint64_t idxCap, uint64_t* total);
/* Replaces CollectFrameEmitters (Sim/SMLMSimulation.cpp:55-101) for a batch of frames on the device. Frame k (index
 * frameIndex[k]) takes the events bucketIdx[bucketStart[k] .. bucketStart[k+1]) (event order; BucketEventsByFrame)
 * that overlap it; driftPx: 2 nFrames (x, y) pixel offsets or NULL; nz > 0: the kernel's plane count and zStepNm for
 * NearestZIndex (0: zIndex 0); illumination: width x height factors at the undrifted site, or NULL; kernel: the
 * records' kernel index. Writes the records into out (outCap records; frame k's are [outStart[k], outStart[k+1]),
 * nFrames + 1 offsets, uint32 as iscg_splat_render_frames takes them), the same bits as the reference's, and zCounts
 * (may be NULL) = {clamped, total}. Arrays may be host or device memory. The call reads the bucket offsets on the
 * host; when out and outStart are device memory, out holds every bucket entry (outCap >= bucketStart[nFrames]) and
 * zCounts is NULL, it then returns without waiting for the device (outStart[nFrames] holds the count), else it
 * synchronizes the context's stream. On a stream being captured into a HIP graph, device bucket offsets must be the
 * array the last eager call on the context validated (same address, same nFrames, unchanged since): its total is
 * reused and every bucket range is clamped to it; such a call can then be captured. */
int32_t iscg_collect_frame_emitters(iscg_context* h, const iscg_blink_event* events, uint32_t nEvents,
                                    const uint32_t* bucketStart, const uint32_t* bucketIdx, const int64_t* frameIndex,
                                    uint32_t nFrames, uint32_t width, uint32_t height, double pixelSizeNm,
                                    double photonsPerBlink, const double* driftPx, int32_t nz, double zStepNm,
                                    double globalZOffsetUm, int32_t kernel, const float* illumination,
                                    iscg_frame_emitter* out, uint64_t outCap, uint32_t* outStart, int64_t* zCounts);

/* Replaces SplatGpuHost::RenderFrames / GpuSimulator::RenderFrames (Sim/SplatGpu.h:57, Sim/GpuSimD3D11.h:99) and,
 * bit for bit, RenderPhotonImage (Sim/SMLMSimulation.cpp:103) + ApplyNoiseChain per frame.
 * ems: nEmitters records, frame k's are [emStart[k], emStart[k+1]) in event order (nFrames + 1 offsets).
 * frames: the frame ids of the noise stream (seed, frames[k], pixel). bgScales: per-frame background factor or NULL
 * (1). extras: nFrames x width*height photons the blinks are added onto instead of the background, or NULL.
 * minPhotons: blinks below it are not splatted (blink-min-photons; 0: all). outAdu (NULL: no noise) and outPhotons
 * (the frame before noise, NULL: not written): nFrames x width*height. */
int32_t iscg_splat_render_frames(iscg_context* h, const iscg_frame_emitter* ems, uint32_t nEmitters,
                                 const uint32_t* emStart, uint32_t nFrames, const uint32_t* frames,
                                 const float* bgScales, const float* extras, const iscg_camera* cam, uint32_t seed,
                                 double minPhotons, uint16_t* outAdu, float* outPhotons);

/* The same for GpuSplatEmitter records (the reference's GPU wire format). */
int32_t iscg_splat_render_records(iscg_context* h, const iscg_splat_emitter* ems, uint32_t nEmitters,
                                  const uint32_t* emStart, uint32_t nFrames, const uint32_t* frames,
                                  const float* bgScales, const float* extras, const iscg_camera* cam, uint32_t seed,
                                  uint16_t* outAdu, float* outPhotons);

/* ---- the binned blink regime (sim::BinnedBlinkRenderer, Sim/BinnedBlinks.h:37-52) ---------------------------------- */

/* Replaces BinnedBlinkRenderer::Setup (Sim/BinnedBlinks.h:43; .cpp:94-163). kernel: the blink kernel's planes (NULL:
 * the in-focus Gaussian of sigmaPx); haloRadius2: the largest radius2 of its halo cut (PsfKernelCache::halo), -1 for
 * none; upscale: requested cells per camera pixel (made a divisor of the oversampling). The grid (W u + 2R) x
 * (H u + 2R) must fit a 1024 x 1024 FFT. Per-plane kernel spectra are made here and kept: a context holds one binned
 * renderer (a second one takes a second context). */
int32_t iscg_binned_setup(iscg_context* h, const iscg_kernel_view* kernel, int32_t haloRadius2, double sigmaPx,
                          uint32_t width, uint32_t height, int32_t upscale);
/* BinnedBlinkRenderer::Upscale / RadiusCells (Sim/BinnedBlinks.h:45-46) of the last setup. */
int32_t iscg_binned_geometry(iscg_context* h, int32_t* upscale, int32_t* radiusCells);
/* Replaces BinnedBlinkRenderer::Render (Sim/BinnedBlinks.h:52; .cpp:165-246) for nFrames frames at once: frame f's
 * emitters are em[frameStart[f] .. frameStart[f+1]) (xPx, yPx, photons, zIndex of sim::FrameEmitter; kernel
 * ignored), in host memory; img (nFrames x height x width, host or device) is added to. */
int32_t iscg_binned_render(iscg_context* h, const iscg_frame_emitter* em, const int32_t* frameStart, int32_t nFrames,
                           float* img);

/* ---- PSF planes ------------------------------------------------------------------------------------------------ */

/* The numeric fields of sim::PsfGeneratorRequest (Sim/PsfGeneratorBridge.h:91-170) for the GibsonLanniZernike model;
 * the Zernike coefficients as the 28 OSA/ANSI values in waves instead of the formatted string. */
typedef struct
{
   double wavelengthNm;
   double na;
   double immersionIndex;
   double sampleIndex;
   double workingDistanceUm;
   double sampleDepthNm;
   double pixelSizeNm;
   double zStepNm;
   double maskWaist;
   double zernike[28];
   int32_t maskType; /* PsfMaskType: 0 none, 1 double helix */
   int32_t maskModes;
   int32_t oversampling;
   int32_t kernelHalfWidthPx;
   int32_t nz;
   int32_t pupilSamples; /* 0: ZernikePupilSamples's choice */
   int32_t interpMode;
   int32_t reserved;
} iscg_psf_request;

/* GeometryFor + ZernikePupilSamples (Sim/ZernikePsf.cpp:118-148): plane size 2 hw os + 1, nz (odd, >= 3), pupil
 * samples per axis. Host only. */
int32_t iscg_psf_geometry(const iscg_psf_request* r, int32_t* size, int32_t* nz, int32_t* pupilSamples);

/* Replaces ComputeZernikePsfPlanes (Sim/ZernikePsf.h:51; normalize != 0: the sum-1 planes of
 * BuildZernikePsfKernelCache, Sim/ZernikePsf.cpp:382-417) for nReq requests sharing size, nz and pupil samples
 * (<= 256). out: nReq x nz x size x size floats (outCap floats available); plane k at defocus (k - (nz-1)/2) zStep. */
int32_t iscg_psf_planes(iscg_context* h, const iscg_psf_request* reqs, int32_t nReq, int32_t normalize, float* out,
                        size_t outCap);

/* A PSF library held on the device, for batches computed again and again (ComputeZernikePsfPlanes over a fixed set
 * of requests; the north star's PSF library of wavelengths x aberration sets). create validates the requests (one
 * size, plane count and pupil sampling, as iscg_psf_planes) and uploads them; run enqueues only kernels (pupil
 * tables, DFT matrices, planes, normalisation) writing out as iscg_psf_planes does, so a run on a device output can
 * be captured into a HIP graph and replayed (one graph launch per library); update replaces the requests' values
 * (same count and geometry, same device buffers: a captured run computes the new requests on its next replay).
 * No reference counterpart as an object: each run equals iscg_psf_planes of the same requests. */
typedef struct iscg_psf_library iscg_psf_library;
int32_t iscg_psf_library_create(iscg_context* h, const iscg_psf_request* reqs, int32_t nReq, iscg_psf_library** lib);
int32_t iscg_psf_library_update(iscg_context* h, iscg_psf_library* lib, const iscg_psf_request* reqs, int32_t nReq);
int32_t iscg_psf_library_geometry(const iscg_psf_library* lib, int32_t* nReq, int32_t* size, int32_t* nz);
int32_t iscg_psf_library_run(iscg_context* h, iscg_psf_library* lib, int32_t normalize, float* out, size_t outCap);
int32_t iscg_psf_library_destroy(iscg_context* h, iscg_psf_library* lib);

/* The kernel half-width the CLI asks for (PsfKernelHalfWidthPx, Sim/PsfGeneratorBridge.cpp:1000): the requested
 * half-width in pixels, raised to 3 Rayleigh radii (capped at 48 px, at least 2). */
int32_t iscg_psf_kernel_half_width_px(double halfWidthNm, double pixelSizeNm, double wavelengthNm, double na);
/* CappedPsfOversampling (Sim/PsfGeneratorBridge.cpp:1011): at most ceil(pixel / maxSample) samples per pixel. */
int32_t iscg_capped_psf_oversampling(int32_t oversampling, double pixelSizeNm, double maxSampleNm);

/* ---- mean field: all foci of a z stack by one 3D FFT convolution ----------------------------------------------- */

/* The geometry of a dye grid and its focus series (WidefieldGridSpec / WidefieldDyeGrid, Sim/WidefieldRender.h:87-115;
 * WidefieldScene::FocusSeries, Sim/WidefieldRender.h:588). Dye plane i is centred at (k0 + i + 0.5) dzUm; image f is
 * at focus focus0Um + f dzUm. The images are the grid cells [fovX0, fovX0 + cw) x [fovY0, fovY0 + ch) (0: to the grid
 * edge). NX, NY, NZ: FFT sizes (powers of two; 0: the smallest with NX >= nx + R, NY >= ny + R, NZ >= nzD + nF - 1,
 * the linear-convolution bound). upscale: grid cells per camera pixel (1 only). */
typedef struct
{
   double dzUm;
   double focus0Um;
   int64_t k0;
   uint32_t nx, ny, nzD, nF;
   uint32_t fovX0, fovY0, cw, ch;
   uint32_t NX, NY, NZ;
   int32_t upscale;
} iscg_wf_desc;

/* Plans the convolution for one PSF (replaces the kernel side of WidefieldScene: KernelWidefieldPsf,
 * Sim/WidefieldRender.cpp:263-323; MakeKernel :1408; PlanFocus's blend and end-plane clamp :1460-1524).
 * psfPlanes: nzP x size x size sum-1 oversampled planes (a PsfKernelCache's), zStepNm apart, plane (nzP-1)/2 in
 * focus. kernelRadius: grid cells (0: KernelWidefieldPsf::Radius). Keeps the kernel spectrum resident. */
int32_t iscg_wf_conv3d_plan(iscg_context* h, const iscg_wf_desc* desc, const float* psfPlanes, int32_t nzP,
                            int32_t size, int32_t oversampling, double zStepNm, int32_t kernelRadius);

/* The planned focus series of one density (replaces WidefieldScene::FocusSeries + WidefieldImages::Render +
 * ApplyNoiseChain for a persistent population, Sim/WidefieldRender.cpp:547-580,2845-2900).
 * density: nzD x ny x nx weighted dye counts. Frame f = max(0, image_f) + background; outImages (nF x ch x cw floats,
 * or NULL) receives it; outAdu (or NULL) its camera ADU, noise stream (seed, frameBase + f, pixel + pixelBase),
 * offset map (ch x cw or NULL: 0). */
int32_t iscg_wf_conv3d(iscg_context* h, const float* density, const iscg_camera* cam, const float* offset,
                       double background, uint32_t seed, uint32_t frameBase, uint32_t pixelBase, float* outImages,
                       uint16_t* outAdu);

/* ---- mean field, one focus per job: the WidefieldAccelerator seam (Sim/WidefieldRender.h:306-361) ---------- */

/* WidefieldGpuJob::Dep (Sim/WidefieldRender.h:325-331): the channel's spectrum gains the plane's times
 * (w0 K[p0] + w1 K[p1]) (p1, w1 only when two). */
typedef struct
{
   uint64_t key;
   int32_t p0, p1;
   float w0, w1;
   int32_t two;
   int32_t reserved;
} iscg_wf_dep;

/* WidefieldGpuJob (Sim/WidefieldRender.h:306-340) flattened into arrays. The index arrays (kernelP, kernelSpec,
 * planeKey, planeStart, channelStart, deps) are host memory; the spectra each kernelSpec entry points to, cells and
 * values may be host or device memory. A plane's cells are distinct (the scene's sparse planes are). */
typedef struct
{
   uint32_t NX, NY;               /* FFT grid: powers of two in [16, 2048] (WidefieldGpuJob::kMaxGpuFft) */
   uint32_t nx, ny;               /* dye grid (<= NX, NY); cell = y * nx + x */
   uint32_t fovX0, fovY0, cw, ch; /* the FOV cells in the grid (fovX0 + cw <= NX, fovY0 + ch <= NY) */
   double fracX, fracY;           /* sub-cell shift */
   uint64_t geometry;             /* keys and kernel planes are valid within one geometry value */
   int32_t nKernels, nPlanes, nChannels, hasPersistent;
   const int32_t* kernelP;        /* [nKernels] kernel plane p */
   const float* const* kernelSpec; /* [nKernels] half spectra: NY rows of NX/2 + 1 complex (re, im) floats */
   const uint64_t* planeKey;      /* [nPlanes] the planes to transform and keep resident */
   const uint32_t* planeStart;    /* [nPlanes + 1] offsets into cells / values, planeStart[0] = 0 */
   const uint32_t* cells;
   const float* values;           /* count x weight */
   const uint32_t* channelStart;  /* [nChannels + 1] offsets into deps */
   const iscg_wf_dep* deps;
} iscg_wf_job;

/* WidefieldAccelerator::HasPlane (Sim/WidefieldRender.h:349): 1 when the spectrum of key in geometry is resident
 * (its dyes are then left out of the job), 0 when not, -1 on error. */
int32_t iscg_wf_has_plane(iscg_context* h, uint64_t geometry, uint64_t key);

/* WidefieldAccelerator::Fits (Sim/WidefieldRender.h:355): 1 when `planes` plane spectra of an NX x NY grid fit in
 * the plane budget beside the resident spectra of the other geometries the context serves, else 0 (the scene then
 * renders on the CPU); -1 on error. A 1 counts the geometry as served. */
int32_t iscg_wf_fits(iscg_context* h, uint64_t geometry, uint32_t NX, uint32_t NY, uint64_t planes);

/* WidefieldAccelerator::Images (Sim/WidefieldRender.h:351; the algorithm of the GPU hosts, restated on the CPU as
 * JobReferenceHost, cli/widefield_check.cpp:816-902): the job's planes are transformed and kept resident by key
 * (kernel spectra by p, within the geometry); per channel, in dep order, S += A_key (w0 K_p0 + w1 K_p1) in complex
 * FP32; the sub-cell phase ramp exp(2 pi i (kx fracX / NX + ky fracY / NY)) (cosine only at the Nyquist bins);
 * the inverse transform, normalised, cropped to the FOV. out: nChannels x ch x cw floats. A dep whose plane or
 * kernel is neither carried nor resident fails the call, as the reference's host does. */
int32_t iscg_wf_images(iscg_context* h, const iscg_wf_job* job, float* out);

/* ---- drifting samples on the mean field: RenderShiftedImages (Sim/WidefieldRender.h:292-300) --------------- */

/* The kept spectra a drift's frames are made from (WidefieldImages with WidefieldScene::SetKeepSpectra,
 * Sim/WidefieldRender.h:281-290): nImages focus images (WidefieldDriftFrames' focus grid), each nChannels channels
 * (the persistent one first when hasPersistent, then the bleach basis), each a RealFft2d half spectrum of NY rows of
 * NX/2 + 1 complex floats; the FOV's cells in the grid, the bin factor and the images' sub-cell pose. */
typedef struct
{
   uint32_t NX, NY;               /* powers of two in [16, 2048] */
   uint32_t fovX0, fovY0, cw, ch; /* FOV cells (cw, ch multiples of upscale) */
   int32_t upscale;
   int32_t nImages, nChannels, hasPersistent;
   double fracX, fracY;           /* WidefieldImages::fracX / fracY */
} iscg_wf_shift_desc;

/* One frame: images i0 and i1 (-1: none) interpolated as (1 - w) i0 + w i1 (w = 0: i0 alone), the sample moved by
 * (dxCells, dyCells) grid cells (WidefieldDriftFrames::Render, Sim/WidefieldRender.cpp:2964-2977). */
typedef struct
{
   int32_t i0, i1;
   double w, dxCells, dyCells;
} iscg_wf_shift_frame;

/* Replaces RenderShiftedImages (Sim/WidefieldRender.cpp:2905-2935) for nFrames frames: frame f's spectrum is the
 * float-coefficient sum of its images' channels (persistent x the image's weight, bleach j x a[f][j] x the weight),
 * phase-ramped by (fracX - dxCells, fracY - dyCells) (ApplyShiftRamp, Sim/DriftShift.cpp:24-48), inverse-transformed
 * on the FOV, max(0, .) and summed over upscale x upscale cells (ShiftedCamera, :50-73), added to cam
 * (nFrames x (ch / upscale) x (cw / upscale) floats). spectra: nImages x nChannels x NY x (NX/2 + 1) complex floats;
 * a: nFrames x (nChannels - hasPersistent) doubles (NULL: no bleach terms). Host or device arrays. */
int32_t iscg_wf_shifted_frames(iscg_context* h, const iscg_wf_shift_desc* desc, const float* spectra,
                               const iscg_wf_shift_frame* frames, const double* a, uint32_t nFrames, float* cam);

/* WidefieldImages::Render (Sim/WidefieldRender.cpp:547-580) for nFrames sets of bleach coefficients, and the camera
 * noise of the frames. channels: [nChannels][ch][cw] cell images (channel 0 the persistent one when hasPersistent,
 * then the bleach maps: iscg_wf_images' output is this layout); a: [nFrames][nChannels - hasPersistent] doubles.
 * Frame k, per camera pixel of the (ch/u) x (cw/u) image: base (base[k], or 0 when base is NULL) + the FP64 sum over
 * its u x u cells of max(0, P + sum_j (float)a[k][j] B_j) (FP32, j up to the frame's last non-zero coefficient), in
 * the reference's operation order (bit for bit). outPhotons: [nFrames][ch/u][cw/u] floats, or NULL. With outAdu: the
 * noise chain of ApplyNoiseChain on those photons (cam, frame ids frames[k], pixel + pixelBase, offset map or NULL),
 * [nFrames][ch/u][cw/u] uint16. */
int32_t iscg_wf_render(iscg_context* h, const float* channels, int32_t nChannels, int32_t hasPersistent, uint32_t cw,
                       uint32_t ch, int32_t upscale, const double* a, uint32_t nFrames, const float* base,
                       const iscg_camera* cam, const float* offset, uint32_t seed, const uint32_t* frames,
                       uint32_t pixelBase, float* outPhotons, uint16_t* outAdu);

/* The device memory the resident plane and kernel spectra of all geometries may take (default: half of the
 * device's). When a job needs more, the least recently used other geometries are dropped first. No reference
 * counterpart (the reference's hosts size their pools themselves). */
int32_t iscg_wf_set_budget(iscg_context* h, uint64_t bytes);

/* ---- differentiable ROI model (new: no reference counterpart) ---------------------------------------------------- */

/* The optics of the ROIs (the fields of a PsfGeneratorRequest that the pupil uses) and the ROI sampling. */
typedef struct
{
   double wavelengthNm, na, immersionIndex, sampleIndex, sampleDepthNm, pixelSizeNm;
   int32_t pupilSamples; /* M, even, <= 64 */
   int32_t roiSize;      /* R pixels per axis; roiSize * oversampling <= 64 */
   int32_t oversampling; /* sub-samples per pixel per axis (the pixel integral) */
   int32_t reserved;
} iscg_roi_model;

/* One emitter: x, y in pixels from the ROI corner (pixel p spans [p, p + 1)), z in nm of defocus (the PSF stack's
 * plane coordinate: ti - ti0 = z - focal shift, Sim/ZernikePsf.cpp:238-244), photons, background photons/pixel. */
typedef struct
{
   double x, y, z, photons, background;
} iscg_roi_params;

/* The expected image mu (nRoi x R x R) of each ROI: the reference's Gibson-Lanni + Zernike pupil (zernike: 28 OSA
 * values in waves) shifted to (x, y), at defocus z, integrated over the pixels, times photons, plus background;
 * the PSF integrates to 1 over the plane. No reference counterpart (the nearest is the finite-difference Fisher
 * matrix of DescribePsfCramerRao, Sim/PsfGeneratorBridge.cpp:1554). */
int32_t iscg_roi_forward(iscg_context* h, const iscg_roi_model* model, const double* zernike,
                         const iscg_roi_params* params, int32_t nRoi, float* mu);

/* The Poisson negative log-likelihood sum (mu - d log mu) of data d (nRoi x R x R counts) and its gradient: per ROI
 * d/d(x, y, z, photons, background) (nRoi x 5) and the shared d/d zernike[28], by the adjoint (pupil recomputed). */
int32_t iscg_roi_backward(iscg_context* h, const iscg_roi_model* model, const double* zernike,
                          const iscg_roi_params* params, int32_t nRoi, const float* data, double* gradParams,
                          double* gradZernike, double* nll);

/* ---- dye schedules --------------------------------------------------------------------------------------------- */

/* One dye's address: the hash H1 of its structure block (DyeH1, core/src/dyes.cpp:175) and its site (k, n). */
typedef struct
{
   uint32_t h1;
   int32_t k;
   int32_t n;
} iscg_dye_address;

/* isc::Kinetics (core/src/dyes.h:129-136). */
typedef struct
{
   double activationRatePerSec, onSec, offSec, bleachProb, photonCV, initialOnSec;
} iscg_kinetics;

/* isc::KineticsHistory (core/src/dyes.h:138-149): n segments of constant kinetics, segment i from tStart[i] (tStart[0]
 * = 0, then strictly increasing, all finite) to tStart[i+1], the last for ever (host arrays of n, at most 65536). The
 * schedule calls take it as LabelSchedule and PersistentBlinks take theirs: NULL or n = 0 none (the kinetics argument
 * from t = 0), n = 1 kin[0] from t = 0, n >= 2 the draws spent over the segments. A call given two or more segments
 * synchronizes the context's stream once, after copying them to the device. */
typedef struct
{
   int32_t n, reserved;
   const double* tStart;
   const iscg_kinetics* kin;
} iscg_kinetics_history;

/* isc::Blink (core/src/dyes.h:197) and isc::BinBlink (:200). */
typedef struct
{
   double tOn, tOff, brightness;
} iscg_blink;
typedef struct
{
   double tOn, tOff, brightness;
   uint32_t bin, j;
} iscg_bin_blink;

/* isc::ContWindow (core/src/dyes.h:190): state 1 PRE, 2 INITIAL_ON, 3 ALWAYS_ON. */
typedef struct
{
   double tOn, tOff;
   uint8_t state;
   double aux;
} iscg_cont_window;

/* The schedule calls below share one output form. dyes: nDyes addresses (host or device). offsets (device or host,
 * nDyes + 1): dye i's records are out[offsets[i] .. offsets[i+1]), offsets[nDyes] = *total, the number of records.
 * out (device or host, capacity records; may be null with capacity 0) receives every dye whose records end at or
 * before capacity; when *total > capacity the call succeeds with the later dyes left out, and a second call with
 * capacity >= *total (the values are a pure function of the addresses) gives all of them. */

/* Replaces isc::LabelSchedule's blinks (core/src/dyes.cpp:440-471, dyes.h:235) for nDyes dyes of one label (mode:
 * 0 dSTORM, 1 PALM, 2 DNA-PAINT, 3 WideField = isc::LabelMode; kin, or hist as above): each dye's blinks
 * in time order, stopping at the first starting at or after tMax, bit-identical to the reference's. */
int32_t iscg_label_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode,
                          const iscg_kinetics* kin, const iscg_kinetics_history* hist, double tMax, int64_t* offsets,
                          iscg_blink* out, int64_t capacity, int64_t* total);

/* Replaces isc::LabelSchedule's continuous windows (core/src/dyes.cpp:447-477): at most one per dye; has[i] = 1 and
 * out[i] set when dye i has one, else has[i] = 0 (out[i] untouched). */
int32_t iscg_label_cont(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, int32_t mode, int32_t preState,
                        const iscg_kinetics* kin, const iscg_kinetics_history* hist, iscg_cont_window* out,
                        uint8_t* has);

/* Replaces isc::PersistentBlinks (core/src/dyes.cpp:636-646, dyes.h:250): each persistent (DNA-PAINT) dye's blinks
 * overlapping the window (t0, t1), by time bin then index (under a history, a bin's later segments' blinks after its
 * own, numbered j | i << 24). Both persistent calls take at most 2^20 time bins per dye and an activation rate of at
 * most 1e5 per bin in every segment (else -1). */
int32_t iscg_persistent_blinks(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                               const iscg_kinetics_history* hist, double t0, double t1, int64_t* offsets,
                               iscg_blink* out, int64_t capacity, int64_t* total);

/* Replaces isc::PersistentBlinksInBins (core/src/dyes.cpp:648-655, dyes.h:256): every blink of bins [binLo, binHi]
 * with its bin and index, unfiltered. */
int32_t iscg_persistent_bins(iscg_context* h, const iscg_dye_address* dyes, int32_t nDyes, const iscg_kinetics* kin,
                             const iscg_kinetics_history* hist, int64_t binLo, int64_t binHi, int64_t* offsets,
                             iscg_bin_blink* out, int64_t capacity, int64_t* total);

/* The device's jsm::sin (fn 0), cos (1), exp (2), log (3), cbrt (4) (core/src/jsmath.h:15-31; fdlibm as V8 runs
 * it) of n values, and V8's Math.hypot of n pairs (fn 5: x holds 2n values) or triples (fn 6: 3n values), the
 * functions the schedules and dye sites use; sin/cos for |x| <= 2^19 pi/2 (NaN beyond). */
int32_t iscg_jsm(iscg_context* h, int32_t fn, const double* x, int64_t n, double* out);

/* ---- the cell field's dye blocks and window queries --------------------------------------------------------------
 * Replaces the block walk of isc::World (core/src/world.cpp:455-640) and the dye generation and queries it drives
 * (GenerateDyes / GenerateNpcDyes / GenerateLaminDyes :808-895 over DyesInBlock / NpcDyesInBlock / LaminDyesInBlock,
 * core/src/dyes.cpp:127-253; SitesInWindow :1212, EventsInWindow :1219, ContinuousInWindow :1279, Density3dInWindow
 * :1519), for one set of kinetics or one kinetics history per structure (no clock map). The cells' geometry is the
 * caller's (the reference's CellAssets: packed cells, microtubule centrelines, pores, lamina points); the device keeps
 * it, builds the microtubule frames (BuildMtFrames, core/src/microtubules.cpp:701) and the dye blocks' reach data
 * once, and each query selects the blocks that reach the window, generates their dyes on the reference's 0.1 nm
 * lattice (World::DyeStore) and answers in the reference's order, bit for bit. */

/* A cell of the field (isc::Cell, core/src/cells.h:48-75: the fields the walk reads). */
typedef struct
{
   int32_t cx, cy;
   double x, y, packRot, rOuter;
} iscg_world_cell;

/* A nuclear pore (isc::Npc, core/src/npcs.h:36-40): cell-local centre, outward normal, ring-plane axes (um), its
 * candidate index. */
typedef struct
{
   double c[3], n[3], e1[3], e2[3];
   int32_t cand, reserved;
} iscg_npc;

/* A point of the nuclear lamina (isc::LaminPoint, core/src/npcs.h:44-48). */
typedef struct
{
   double c[3], n[3];
   int32_t cand, reserved;
} iscg_lamin_point;

/* The cells' geometry (host arrays; isc::CellAssets, core/src/world.h:58-90):
 *   cells[nCells] in the walk's order (World::CellsInRectPtr: pack block x, then y, then the block's order);
 *   cell c's microtubules are [mtStart[c], mtStart[c+1]) (nCells + 1 entries); microtubule m's centreline is the
 *   points pts[3 ptStart[m] .. 3 ptStart[m+1]) (x, y, z, cell-local um) and mtIndex[m] its index in its cell (null:
 *   its position in the cell's range);
 *   cell c's pores npcs[npcStart[c] .. npcStart[c+1]) and lamina points lamin[laminStart[c] .. laminStart[c+1]), in
 *   candidate order (null start arrays: none);
 *   npcSites: NpcAttachmentPoints of the params (64 x 3, nm); npcSiteReachUm: NpcSiteReachUm. */
typedef struct
{
   int32_t nCells;
   const iscg_world_cell* cells;
   const int32_t* mtStart;
   const int64_t* ptStart;
   const int32_t* mtIndex;
   const double* pts;
   const int32_t* npcStart;
   const iscg_npc* npcs;
   const int32_t* laminStart;
   const iscg_lamin_point* lamin;
   double npcSites[192];
   double npcSiteReachUm;
} iscg_world_geometry;

/* A structure's label (isc::Label, core/src/dyes.h:163-176): which sites carry a dye and how it emits. mode 0
 * dSTORM, 1 PALM, 2 DNA-PAINT, 3 WideField; preState PALM only; hasLinkage 0 takes the structure's built-in linkage
 * (LegacyLinkage), and the linkage is normalised (NormLinkage) as World::SetLabel does. The orientation does not
 * change the dyes or their emission and is not taken. */
typedef struct
{
   double density, fluorescentFraction;
   int32_t mode, preState;
   iscg_kinetics kin;
   int32_t hasLinkage, reserved;
   double binderNm, binderFlexNm, linkerMinNm, linkerMaxNm;
} iscg_label;

/* A dye of SitesInWindow (isc::WorldDye, core/src/world.h:92-99). */
typedef struct
{
   double x, y, z;
   uint32_t id;
   int32_t cx, cy, mtIndex, k, n, structure, reserved;
} iscg_world_dye;

/* A blink or continuous window (isc::WorldEvent, core/src/world.h:103-111, without its cache key). */
typedef struct
{
   double x, y, z, tOn, tOff, brightness, aux;
   uint32_t id;
   uint8_t structure, state, reserved[2];
} iscg_world_event;

/* A dye of one walked block as the generators make it (isc::Dye / isc::NpcDye, core/src/dyes.h:263-300): the
 * cell-local, unrounded position (before World::GenerateDyes' LocalToWorld and lattice), its id, the block's place
 * in the walk, the structure, the element (microtubule index in its cell; a pore's or lamina point's index in its
 * cell's list) and the site (k, n). */
typedef struct
{
   double x, y, z;
   uint32_t id;
   int32_t block, structure, element, k, n, reserved;
} iscg_block_dye;

typedef struct iscg_world iscg_world;

/* Keeps the geometry on the context's device (isc::World over its cell assets, core/src/world.h:152): the default
 * labels are the reference's (density 0.7, every dye fluorescent, DNA-PAINT, isc::Kinetics' defaults, the built-in
 * linkages). The world uses h's stream and reports errors on h; destroy it before h. */
int32_t iscg_world_create(iscg_context* h, uint32_t seed, const iscg_world_geometry* geometry, iscg_world** out);
int32_t iscg_world_destroy(iscg_world* w);
/* World::SetLabel (core/src/world.cpp:1080) for structure 0 (microtubules), 1 (pores), 2 (lamina) or 3 (docking
 * strands: no geometry here); -1 for an invalid label (as ValidateLabel, kinetics as the C ABI checks them). */
int32_t iscg_world_set_label(iscg_world* w, int32_t structure, const iscg_label* label);
/* World::SetKineticsHistory (core/src/world.cpp:933-977): n segments starting at tStart[0..n) (0, then finite and
 * increasing), structure s's kinetics in segment i at kin[i * 4 + s]; n = 0 clears it. The events and continuous
 * windows of a structure then follow its history (HistoryOf) as LabelSchedule and PersistentBlinksInBins do under
 * it; a label's own kinetics apply again once the history is cleared. -1 (nothing changes) on bad input. */
int32_t iscg_world_set_history(iscg_world* w, const double* tStart, const iscg_kinetics* kin, int32_t n);

/* The queries take the window as {x0, y0, x1, y1, zMin, zMax} (world um) and write their first min(*total, capacity)
 * rows (host or device out; *total receives the count). */
/* DyesInBlock / NpcDyesInBlock / LaminDyesInBlock (core/src/dyes.cpp:127-253) for every block the window's walk
 * takes (World::ForEachDyeBlock's selection, core/src/world.cpp:556-589), blocks in walk order, each block's dyes in
 * the generator's order, not filtered by the window. */
int32_t iscg_world_block_dyes(iscg_world* w, const double window[6], iscg_block_dye* out, int64_t capacity,
                              int64_t* total);
/* World::SitesInWindow (core/src/world.cpp:1212). */
int32_t iscg_world_sites(iscg_world* w, const double window[6], iscg_world_dye* out, int64_t capacity, int64_t* total);
/* World::EventsInWindow (core/src/world.cpp:1219): the blinks overlapping [t0, t1), per dye block in walk order:
 * dSTORM / PALM blinks by tOn (ties in dye and blink order), DNA-PAINT blinks in (dye, bin, index) order. */
int32_t iscg_world_events(iscg_world* w, const double window[6], double t0, double t1, iscg_world_event* out,
                          int64_t capacity, int64_t* total);
/* World::ContinuousInWindow (core/src/world.cpp:1279): the continuous windows ending after tMin, in dye order. */
int32_t iscg_world_continuous(iscg_world* w, const double window[6], double tMin, iscg_world_event* out,
                              int64_t capacity, int64_t* total);
/* World::Density3dInWindow (core/src/world.cpp:1519): counts on nx x ny x nz voxels, out[(k ny + iy) nx + ix] (host
 * or device), structures of structureMask; *total receives the count. */
int32_t iscg_world_density3d(iscg_world* w, const double window[6], int32_t nx, int32_t ny, int32_t nz,
                             uint32_t structureMask, float* out, int64_t* total);

#ifdef __cplusplus
}
#endif

#endif
