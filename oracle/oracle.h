/*
 * oracle.h -- FP64 CPU restatement of inSiliScope's image-formation path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library (insiliscope_amd). Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository; Sim/ = adapter/inSiliScope/Simulation/).
 *
 * Parity status of each part (how it is pinned) is listed at the top of
 * oracle.c.
 */
#ifndef ISC_ORACLE_H
#define ISC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- address / counter RNG (core/src/rng.h:18-52, Sim/SMLMCounterRng.h:37-140) ---- */
void orc_pcg4d(uint32_t v[4]);
double orc_hash_unit(uint32_t seed, int32_t cx, int32_t cy, uint32_t k);
void orc_hash_stream(uint32_t seed, int32_t cx, int32_t cy, uint32_t base, int32_t n, double* out);
/* n uniforms of the counter stream (seed, frame, pixel), counter from 0. */
void orc_counter_uniforms(uint32_t seed, uint32_t frame, uint32_t pixel, int32_t n, double* out);
/* n draws of CounterGauss / CounterPoisson(l) / CounterGamma(k) on a fresh pixel stream. */
void orc_counter_gauss(uint32_t seed, uint32_t frame, uint32_t pixel, int32_t n, double* out);
double orc_counter_poisson(uint32_t seed, uint32_t frame, uint32_t pixel, double l);
double orc_counter_gamma(uint32_t seed, uint32_t frame, uint32_t pixel, double k);

/* ---- camera noise (Sim/SMLMNoise.h:35-50, Sim/SMLMNoise.cpp:142-216) ---- */
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
} orc_camera;

/* offset/gain/read maps: width*height floats or NULL (NULL = the scalar; a NULL offset map means offset 0). */
void orc_noise_chain(const float* photons, uint16_t* out, uint32_t width, uint32_t height, const orc_camera* cam,
                     const float* offsetMap, const float* gainMap, const float* readMap, uint32_t seed,
                     uint32_t frame, uint32_t pixelBase);

/* ---- Gibson-Lanni + Zernike PSF (Sim/ZernikePsf.cpp, Sim/ChirpZ.h, Sim/FftRadix2.h) ---- */
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
   double zernike[28]; /* OSA/ANSI 0-27, waves */
   int32_t maskType;   /* 0 none, 1 double helix */
   int32_t maskModes;
   int32_t oversampling;
   int32_t kernelHalfWidthPx;
   int32_t nz;
   int32_t pupilSamples;
   int32_t interpMode; /* 0 nearest, 1 linear, 2 cubic */
   int32_t reserved;
} orc_psf_request;

/* size (2 hw os + 1), nz (odd, >= 3) and pupil samples M; 0 on success. */
int32_t orc_psf_geometry(const orc_psf_request* r, int32_t* size, int32_t* nz, int32_t* M);
/* Raw |E|^2 planes, nz * size * size floats (ComputeZernikePsfPlanes). 0 on success, -1 on bad request. */
int32_t orc_psf_planes(const orc_psf_request* r, float* out);
/* Sum-1 normalisation of each plane in place (BuildZernikePsfKernelCache). */
void orc_psf_normalize(float* planes, int32_t nz, int32_t n);
/* The same planes as FP64 complex fields before |E|^2 (re, im interleaved): nz * size * size * 2 doubles. */
int32_t orc_psf_fields(const orc_psf_request* r, double* out);

/* ---- kernel cache / splat (Sim/PsfGeneratorBridge.cpp:40-60,1021-1513, Sim/SplatKernel.inl) ---- */
void orc_block_sums(const float* kernel, int32_t n, int32_t os, float* out);
int32_t orc_nearest_z(int32_t nz, double zStepNm, double zUm, int32_t* clamped);

typedef struct
{
   int32_t x0, y0, bx, by, nTaps, pad;
   double wx[4], wy[4];
} orc_splat_setup_t;

void orc_splat_setup(int32_t sizeOversampled, int32_t os, double xPx, double yPx, int32_t interp,
                     orc_splat_setup_t* out);

/* One frame emitter as RenderPhotonImage places it (Sim/SMLMSimulation.h:158). */
typedef struct
{
   double xPx, yPx, photons;
   int32_t zIndex, kernel;
} orc_emitter;

/* A frame: background (flat bg, or bgMap[i]*bgScale when bgMap != NULL), then every emitter in order. blockSums:
 * nz planes of bw*bw row-major. spans: NULL or nz*(2 camRad+1) int16 pairs. */
void orc_render_frame(float* img, uint32_t width, uint32_t height, double bg, const float* bgMap, double bgScale,
                      const orc_emitter* ems, int32_t nEms, const float* blockSums, int32_t nzK, int32_t bw,
                      int32_t os, int32_t sizeOversampled, int32_t interp, const int16_t* spanLo,
                      const int16_t* spanHi, double minPhotons);

/* FftShiftKernelTile (Sim/PsfGeneratorBridge.cpp:1293-1380): the n x n kernel moved by (shiftX, shiftY) samples,
 * g(i) = f(i + shift), by zero-padded FP64 radix-2 Fourier shifts along rows then columns. */
void orc_fft_shift_tile(const float* kernel, int32_t n, double shiftX, double shiftY, float* out);

/* RenderPhotonImage with the Fft placement (PlanSplat, Sim/PsfGeneratorBridge.cpp:1443-1456): per emitter one
 * Fourier shift of its raw plane (planes: nz x n x n) and a nearest read of the shifted block sums. */
void orc_render_frame_fft(float* img, uint32_t width, uint32_t height, double bg, const float* bgMap, double bgScale,
                          const orc_emitter* ems, int32_t nEms, const float* planes, int32_t nzK, int32_t os,
                          int32_t n, const int16_t* spanLo, const int16_t* spanHi, double minPhotons);

/* WidefieldImages::Render (Sim/WidefieldRender.cpp:547-580): channels [nCh][ch][cw] (channel 0 persistent when
 * hasPersistent), bleach coefficients a[nA], added to cam ((ch/u) x (cw/u)). At most 64 bleach channels. */
void orc_wf_images_render(const float* channels, int32_t nCh, int32_t hasPersistent, uint32_t cw, uint32_t ch,
                          int32_t upscale, const double* a, int32_t nA, float* cam);

/* ---- mean field (Sim/WidefieldRender.cpp:263-323,1460-1524; cli/widefield_check.cpp:287-343) ---- */
/* Grid-pitch kernel plane p of an oversampled plane (KernelWidefieldPsf::Kernel): (2R+1)^2 floats. */
int32_t orc_wf_radius(int32_t sizeOversampled, int32_t os, int32_t upscale);
void orc_wf_kernel(const float* plane, int32_t sizeOversampled, int32_t os, int32_t upscale, int32_t R, float* out);
/* Per-focus direct convolution of dye planes (density nzD x ny x nx, already weighted; plane i centred at
 * (k0 + i + 0.5) dz um) with PSF planes (nzP planes of (2R+1)^2), linear z blend with clamp at the end planes as
 * PlanFocus: t = ((k0 + i + 0.5) dz - focus_f) * 1000 / zStepNm + (nzP - 1) / 2. out: nF x ny x nx, FP64, before
 * max(0, .) and binning. */
void orc_wf_plane_weight(double centreUm, double focusUm, int32_t nzP, double zStepNm, int32_t* p0Out,
                         double* frOut);
void orc_wf_direct(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzUm, int64_t k0,
                   const float* kernels, int32_t nzP, int32_t R, double zStepNm, const double* focusUm, int32_t nF,
                   double* out);

/* ---- ROI forward model and Poisson-NLL gradient (new capability; no reference counterpart) ---- */
/* The reference's pupil (Sim/ZernikePsf.cpp:163-248) evaluated at an emitter: E(x_s) = sum_mn P_mn(z)
 * exp(i(k_m (x_s - x) + k_n (y_s - y))) on os x os sub-samples of an R x R pixel ROI, mu = N norm sum_sub |E|^2 + b,
 * norm = (dxs dk / 2 pi)^2 / n_inside (the PSF's integral over one period is 1, Parseval). */
typedef struct
{
   double wavelengthNm, na, immersionIndex, sampleIndex, sampleDepthNm, pixelSizeNm;
   int32_t pupilSamples; /* M */
   int32_t roiSize;      /* R pixels per axis */
   int32_t oversampling; /* sub-samples per pixel per axis */
   int32_t reserved;
} orc_roi_model;

typedef struct
{
   double x, y; /* pixels from the ROI corner (pixel p spans [p, p + 1)) */
   double z;    /* nm of defocus (the PSF stack's plane coordinate) */
   double photons, background;
} orc_roi_params;

/* mu: nRoi x R x R. */
void orc_roi_forward(const orc_roi_model* m, const double* zernike, const orc_roi_params* p, int32_t nRoi,
                     double* mu);
/* NLL = sum (mu - d log mu); gradients d NLL / d (x, y, z, N, b) per ROI (nRoi x 5) and d NLL / d zernike[28]. */
double orc_roi_backward(const orc_roi_model* m, const double* zernike, const orc_roi_params* p, int32_t nRoi,
                        const double* data, double* gradParams, double* gradZernike);

/* the binned blink regime (Sim/BinnedBlinks.cpp) */
void orc_binned_geometry(int32_t hasKernel, int32_t size, int32_t os, int32_t upscale, double sigmaPx, int32_t haloR2,
                         int32_t* u, int32_t* R);
void orc_binned_gauss(double sigmaPx, int32_t u, int32_t R, float* out);
void orc_binned_render(const float* kernels, int32_t nz, int32_t u, int32_t R, int32_t W, int32_t H, const double* em,
                       int32_t n, float* img);

/* ---- JS-compatible FP64 math (core/src/jsmath_fdlibm.cpp: fdlibm 5.3 as V8 runs it) ----
 * sin/cos restate the reductions for |x| <= 2^19 pi/2 only (beyond: NaN). */
double orc_jsm_sin(double x);
double orc_jsm_cos(double x);
double orc_jsm_exp(double x);
double orc_jsm_log(double x);

/* ---- dye schedules (core/src/dyes.h:129-261, core/src/dyes.cpp:254-657), one set of kinetics ---- */
typedef struct
{
   double activationRatePerSec, onSec, offSec, bleachProb, photonCV, initialOnSec;
} orc_kinetics;
/* LabelSchedule's blinks of dye (h1, k, n) under mode (0 dSTORM, 1 PALM, 2 DNA-PAINT, 3 WideField): out holds
 * cap x 3 doubles (tOn, tOff, brightness); returns the count (written up to cap). */
int32_t orc_label_blinks(uint32_t h1, int32_t k, int32_t n, int32_t mode, const orc_kinetics* kin, double tMax,
                         double* out, int32_t cap);
/* LabelSchedule's continuous window (at most one for a single set of kinetics): 1 and w = tOn, tOff, state, aux,
 * or 0. */
int32_t orc_label_cont(uint32_t h1, int32_t k, int32_t n, int32_t mode, int32_t preState, const orc_kinetics* kin,
                       double w[4]);
/* PersistentGen: window (t0, t1) when bins is 0 (PersistentBlinks), else every blink of bins [binLo, binHi]
 * (PersistentBlinksInBins). out: cap x 5 doubles (tOn, tOff, brightness, bin, j); returns the count. */
int32_t orc_persistent_blinks(uint32_t h1, int32_t k, int32_t n, const orc_kinetics* kin, int32_t bins, double t0,
                              double t1, int64_t binLo, int64_t binHi, double* out, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif
