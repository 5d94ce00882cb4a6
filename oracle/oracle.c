/*
 * oracle.c -- FP64 CPU restatement of inSiliScope's image-formation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see oracle.h). It is
 * compiled with -ffp-contract=off -fno-fast-math like the reference
 * (core/CMakeLists.txt:41), and every expression keeps the reference's operand
 * order, so where the reference is deterministic FP64 this file gives the same
 * bits on the same libm.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   pcg4d / HashUnit / HashStream ... spec/golden/ref.txt `rng` and `stream` lines (bit-exact).
 *   counter uniforms/Gauss/Poisson/Gamma, noise chain ... ADU fixtures written by the compiled reference
 *                                     (oracle/_ref) in tests/golden (bit-exact); moments and closed forms.
 *   PSF planes ....................... tests/psf/zernike_ref.bin (4 cases, rel L2 < 1e-6, same argmax); compiled
 *                                     reference planes at the C1/C3 geometries (bit-exact); chirp-Z == direct DFT
 *                                     on tiny grids; sum-1 after normalisation.
 *   block sums / splat ............... compiled-reference frames in tests/golden (bit-exact); photon conservation;
 *                                     block sums total os^2.
 *   mean-field kernel + direct conv .. compiled-reference KernelWidefieldPsf planes (bit-exact); delta density
 *                                     reproduces the kernel; linearity; brute force on tiny grids.
 *   jsm log / exp / sin / cos ........ spec/golden/ref.txt `m` lines (sin/cos for |x| <= 2^19 pi/2), compiled
 *                                     reference at the reduction boundaries (bit-exact); libm to one ulp
 *                                     (tests/test_oracle_dyes.py).
 *   dye schedules .................... compiled-reference LabelSchedule / PersistentBlinks(InBins) fixtures and
 *                                     random kinetics (bit-exact); the activation, ON, brightness, bleach and
 *                                     Poisson laws (tests/test_oracle_dyes.py).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.14159265358979323846;

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ======================================================================= */
/* RNG: core/src/rng.h:18-52, Sim/SMLMCounterRng.h:37-55                  */
/* ======================================================================= */

void orc_pcg4d(uint32_t v[4])
{
   uint32_t a = v[0], b = v[1], c = v[2], d = v[3];
   a = a * 1664525u + 1013904223u;
   b = b * 1664525u + 1013904223u;
   c = c * 1664525u + 1013904223u;
   d = d * 1664525u + 1013904223u;
   a += b * d;
   b += c * a;
   c += a * b;
   d += b * c;
   a ^= a >> 16;
   b ^= b >> 16;
   c ^= c >> 16;
   d ^= d >> 16;
   a += b * d;
   b += c * a;
   c += a * b;
   d += b * c;
   v[0] = a;
   v[1] = b;
   v[2] = c;
   v[3] = d;
}

/* core/src/rng.h:37 */
double orc_hash_unit(uint32_t seed, int32_t cx, int32_t cy, uint32_t k)
{
   uint32_t v[4] = {seed, (uint32_t)cx, (uint32_t)cy, k};
   orc_pcg4d(v);
   return ((double)(v[0] >> 9) + 0.5) * 1.1920928955078125e-7;
}

/* core/src/rng.h:45-52: draw i = HashUnit(seed, cx, cy, base*4096 + i), uint32 wrap. */
void orc_hash_stream(uint32_t seed, int32_t cx, int32_t cy, uint32_t base, int32_t n, double* out)
{
   const uint32_t b = base * 4096u;
   for (int32_t i = 0; i < n; ++i)
      out[i] = orc_hash_unit(seed, cx, cy, b + (uint32_t)i);
}

/* Sim/SMLMCounterRng.h:59-89 (CounterRng). */
typedef struct
{
   uint32_t seed, frame, pix, ctr;
   uint32_t buf[4];
   int k;
} crng;

static void crng_init(crng* u, uint32_t seed, uint32_t frame, uint32_t pixel)
{
   u->seed = seed;
   u->frame = frame;
   u->pix = pixel;
   u->ctr = 0;
   u->k = 4;
}

static double crng_uniform(crng* u)
{
   if (u->k == 4)
   {
      u->buf[0] = u->seed;
      u->buf[1] = u->frame;
      u->buf[2] = u->pix;
      u->buf[3] = u->ctr;
      orc_pcg4d(u->buf);
      ++u->ctr;
      u->k = 0;
   }
   return ((u->buf[u->k++] >> 9) + 0.5) * 1.1920928955078125e-7;
}

/* Sim/SMLMCounterRng.h:95-99 */
static double crng_gauss(crng* u)
{
   double r = sqrt(-2.0 * log(crng_uniform(u)));
   return r * cos(6.283185307179586 * crng_uniform(u));
}

/* Sim/SMLMCounterRng.h:104-119 */
static double crng_poisson(double l, crng* u)
{
   if (l > 60.0)
   {
      double v = floor(l + sqrt(l) * crng_gauss(u) + 0.5);
      return v > 0.0 ? v : 0.0;
   }
   if (!(l > 0.0))
      return 0.0;
   const double L = exp(-l);
   int k = 0;
   double p = 1.0;
   do
   {
      ++k;
      p *= crng_uniform(u);
   } while (p > L && k < 1000);
   return (double)(k - 1);
}

/* Sim/SMLMCounterRng.h:123-138 */
static double crng_gamma(double k, crng* u)
{
   const double d = k - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
   for (int it = 0; it < 1000; ++it)
   {
      const double x = crng_gauss(u), t = 1.0 + c * x;
      if (t <= 0.0)
         continue;
      const double v = t * t * t, w = crng_uniform(u);
      if (w < 1.0 - 0.0331 * x * x * x * x)
         return d * v;
      if (log(w) < 0.5 * x * x + d * (1.0 - v + log(v)))
         return d * v;
   }
   return d;
}

void orc_counter_uniforms(uint32_t seed, uint32_t frame, uint32_t pixel, int32_t n, double* out)
{
   crng u;
   crng_init(&u, seed, frame, pixel);
   for (int32_t i = 0; i < n; ++i)
      out[i] = crng_uniform(&u);
}

void orc_counter_gauss(uint32_t seed, uint32_t frame, uint32_t pixel, int32_t n, double* out)
{
   crng u;
   crng_init(&u, seed, frame, pixel);
   for (int32_t i = 0; i < n; ++i)
      out[i] = crng_gauss(&u);
}

double orc_counter_poisson(uint32_t seed, uint32_t frame, uint32_t pixel, double l)
{
   crng u;
   crng_init(&u, seed, frame, pixel);
   return crng_poisson(l, &u);
}

double orc_counter_gamma(uint32_t seed, uint32_t frame, uint32_t pixel, double k)
{
   crng u;
   crng_init(&u, seed, frame, pixel);
   return crng_gamma(k, &u);
}

/* ======================================================================= */
/* Camera noise: Sim/SMLMNoise.cpp:142-216                                 */
/* ======================================================================= */

void orc_noise_chain(const float* photons, uint16_t* out, uint32_t width, uint32_t height, const orc_camera* cam,
                     const float* offsetMap, const float* gainMap, const float* readMap, uint32_t seed,
                     uint32_t frame, uint32_t pixelBase)
{
   const size_t n = (size_t)width * height;
   crng u;
   crng_init(&u, seed, frame, 0);
   if (cam->emccd)
   {
      const int bits = imin(16, imax(1, cam->bitDepth));
      const double maxAdu = ldexp(1.0, bits) - 1.0;
      const double emGain = cam->emGain > 1.0 ? cam->emGain : 1.0;
      for (size_t i = 0; i < n; ++i)
      {
         crng_init(&u, seed, frame, (uint32_t)i + pixelBase);
         const double ph0 = (double)photons[i];
         const double ph = ph0 > 0.0 ? ph0 : 0.0;
         const double ne =
            crng_poisson(ph * cam->quantumEfficiency + cam->darkCurrentElectrons + cam->cicElectrons, &u);
         const double o = ne > 0.0 ? crng_gamma(ne, &u) : 0.0;
         const double readE = (readMap ? (double)readMap[i] : cam->readNoiseElectrons) / emGain;
         const double gain = gainMap ? (double)gainMap[i] : cam->gainPhotonsPerAdu;
         const double offset = offsetMap ? (double)offsetMap[i] : 0.0;
         const double g = crng_gauss(&u);
         double adu = floor(offset + (o + readE * g) / gain + 0.5);
         out[i] = (uint16_t)(adu < 0.0 ? 0.0 : (adu > maxAdu ? maxAdu : adu));
      }
      return;
   }
   for (size_t i = 0; i < n; ++i)
   {
      crng_init(&u, seed, frame, (uint32_t)i + pixelBase);
      const double ph0 = (double)photons[i];
      const double ph = ph0 > 0.0 ? ph0 : 0.0;
      const double ne = crng_poisson(ph * cam->quantumEfficiency + cam->darkCurrentElectrons, &u);
      const double readNoise = readMap ? (double)readMap[i] : cam->readNoiseElectrons;
      const double gain = gainMap ? (double)gainMap[i] : cam->gainPhotonsPerAdu;
      const double offset = offsetMap ? (double)offsetMap[i] : 0.0;
      const double g = crng_gauss(&u);
      double adu = offset + (ne + readNoise * g) / gain;
      adu = adu < 0.0 ? 0.0 : (adu > 65535.0 ? 65535.0 : adu);
      out[i] = (uint16_t)floor(adu + 0.5);
   }
}

/* ======================================================================= */
/* Radix-2 FFT and chirp-Z: Sim/FftRadix2.h:39-113, Sim/ChirpZ.h:28-141     */
/* ======================================================================= */

typedef struct
{
   int n;
   double *c, *s;
   int* rev;
} twiddles;

/* Sim/FftRadix2.h:39 MakeTwiddles */
static void make_twiddles(twiddles* t, int n, int sign)
{
   t->n = n;
   t->c = (double*)calloc((size_t)imax(1, n - 1), sizeof(double));
   t->s = (double*)calloc((size_t)imax(1, n - 1), sizeof(double));
   t->rev = (int*)calloc((size_t)imax(1, n), sizeof(int));
   for (int len = 2; len <= n; len <<= 1)
   {
      const double ang = sign * 2.0 * 3.14159265358979323846 / len;
      const double wr = cos(ang), wi = sin(ang);
      const int half = len / 2;
      double cr = 1.0, ci = 0.0;
      for (int k = 0; k < half; ++k)
      {
         t->c[half - 1 + k] = cr;
         t->s[half - 1 + k] = ci;
         const double nr = cr * wr - ci * wi;
         ci = cr * wi + ci * wr;
         cr = nr;
      }
   }
   for (int i = 1, j = 0; i < n; ++i)
   {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1)
         j ^= bit;
      j ^= bit;
      t->rev[i] = j;
   }
}

static void free_twiddles(twiddles* t)
{
   free(t->c);
   free(t->s);
   free(t->rev);
}

/* Sim/FftRadix2.h:74 Fft1d */
static void fft1d(double* re, double* im, const twiddles* tw)
{
   const int n = tw->n;
   for (int i = 1; i < n; ++i)
   {
      const int j = tw->rev[i];
      if (i < j)
      {
         double t = re[i];
         re[i] = re[j];
         re[j] = t;
         t = im[i];
         im[i] = im[j];
         im[j] = t;
      }
   }
   for (int len = 2; len <= n; len <<= 1)
   {
      const int half = len / 2;
      const double* tc = tw->c + (half - 1);
      const double* ts = tw->s + (half - 1);
      for (int i = 0; i < n; i += len)
      {
         double* ar = re + i;
         double* ai = im + i;
         double* br = re + i + half;
         double* bi = im + i + half;
         for (int k = 0; k < half; ++k)
         {
            const double cr = tc[k], ci = ts[k];
            const double vr = br[k] * cr - bi[k] * ci, vi = br[k] * ci + bi[k] * cr;
            br[k] = ar[k] - vr;
            bi[k] = ai[k] - vi;
            ar[k] += vr;
            ai[k] += vi;
         }
      }
   }
}

typedef struct
{
   int M, P, L;
   double invL;
   twiddles fwd, inv;
   double *preC, *preS, *gRe, *gIm, *c1, *s1, *c2, *s2;
} czt_plan;

/* Sim/ChirpZ.h:40-84 CztPlan(M, dk, k0, P, dx, x0) */
static void czt_init(czt_plan* p, int M, double dk, double k0, int P, double dx, double x0)
{
   p->M = M;
   p->P = P;
   const double theta = dk * dx;
   int L = 1;
   while (L < M + P - 1)
      L <<= 1;
   p->L = L;
   p->invL = 1.0 / L;
   make_twiddles(&p->fwd, L, -1);
   make_twiddles(&p->inv, L, 1);
   p->preC = (double*)malloc(sizeof(double) * (size_t)M);
   p->preS = (double*)malloc(sizeof(double) * (size_t)M);
   for (int m = 0; m < M; ++m)
   {
      const double ang = m * dk * x0 + theta * m * m / 2.0;
      p->preC[m] = cos(ang);
      p->preS[m] = sin(ang);
   }
   p->gRe = (double*)calloc((size_t)L, sizeof(double));
   p->gIm = (double*)calloc((size_t)L, sizeof(double));
   for (int n = -(M - 1); n < P; ++n)
   {
      const double ang = -theta * n * n / 2.0;
      const int idx = n >= 0 ? n : L + n;
      p->gRe[idx] = cos(ang);
      p->gIm[idx] = sin(ang);
   }
   fft1d(p->gRe, p->gIm, &p->fwd);
   p->c1 = (double*)malloc(sizeof(double) * (size_t)P);
   p->s1 = (double*)malloc(sizeof(double) * (size_t)P);
   p->c2 = (double*)malloc(sizeof(double) * (size_t)P);
   p->s2 = (double*)malloc(sizeof(double) * (size_t)P);
   for (int q = 0; q < P; ++q)
   {
      const double a1 = theta * q * q / 2.0;
      p->c1[q] = cos(a1);
      p->s1[q] = sin(a1);
      const double a2 = k0 * (x0 + q * dx);
      p->c2[q] = cos(a2);
      p->s2[q] = sin(a2);
   }
}

static void czt_free(czt_plan* p)
{
   free_twiddles(&p->fwd);
   free_twiddles(&p->inv);
   free(p->preC);
   free(p->preS);
   free(p->gRe);
   free(p->gIm);
   free(p->c1);
   free(p->s1);
   free(p->c2);
   free(p->s2);
}

/* Sim/ChirpZ.h:91-141 Apply, one line (each batched line performs exactly these operations). The reference's
 * pruned first stages give the same values up to the sign of zeros, which |E|^2 cannot see (Sim/FftRadix2.h:110-118). */
static void czt_apply(const czt_plan* p, const double* inRe, const double* inIm, double* outRe, double* outIm,
                      double* aRe, double* aIm)
{
   const int L = p->L;
   memset(aRe, 0, sizeof(double) * (size_t)L);
   memset(aIm, 0, sizeof(double) * (size_t)L);
   for (int m = 0; m < p->M; ++m)
   {
      const double cr = p->preC[m], ci = p->preS[m];
      const double re = inRe[m], im = inIm[m];
      aRe[m] = re * cr - im * ci;
      aIm[m] = re * ci + im * cr;
   }
   fft1d(aRe, aIm, &p->fwd);
   for (int i = 0; i < L; ++i)
   {
      const double gr = p->gRe[i], gi = p->gIm[i];
      const double re = aRe[i] * gr - aIm[i] * gi;
      const double im = aRe[i] * gi + aIm[i] * gr;
      aRe[i] = re;
      aIm[i] = im;
   }
   fft1d(aRe, aIm, &p->inv);
   for (int q = 0; q < p->P; ++q)
   {
      const double cc1 = p->c1[q], ss1 = p->s1[q];
      const double cc2 = p->c2[q], ss2 = p->s2[q];
      const double convRe = aRe[q] * p->invL, convIm = aIm[q] * p->invL;
      const double sRe = convRe * cc1 - convIm * ss1, sIm = convRe * ss1 + convIm * cc1;
      outRe[q] = sRe * cc2 - sIm * ss2;
      outIm[q] = sRe * ss2 + sIm * cc2;
   }
}

/* ======================================================================= */
/* Zernike basis and pupil: Sim/ZernikePsf.cpp:35-110                      */
/* ======================================================================= */

/* :35 IndexToNM */
static void index_to_nm(int j, int* n, int* m)
{
   for (int nn = 0;; ++nn)
      for (int l = 0; l <= nn; ++l)
         if (nn * (nn + 1) / 2 + l == j)
         {
            *n = nn;
            *m = -nn + 2 * l;
            return;
         }
}

/* :47 Binomial */
static double binomial(int a, int b)
{
   if (b < 0 || b > a)
      return 0.0;
   double r = 1.0;
   for (int i = 0; i < b; ++i)
      r = r * (a - i) / (i + 1);
   return r;
}

/* :58 ZernikeRadial (unnormalised, m >= 0) */
static double zernike_radial(int n, int m, double rho)
{
   double r = 0.0;
   for (int k = 0; k <= (n - m) / 2; ++k)
   {
      const double coeff = (k % 2 == 0 ? 1.0 : -1.0) * binomial(n - k, k) * binomial(n - 2 * k, (n - m) / 2 - k);
      r += coeff * pow(rho, (double)(n - 2 * k));
   }
   return r;
}

/* :74 ZernikeValue */
static double zernike_value(int n, int l, int m, double rho, double phi)
{
   const double radial = rho <= 1.0 ? zernike_radial(n, m, rho) : 0.0;
   return radial * (l >= 0 ? cos(m * phi) : sin(m * phi));
}

/* :80 LaguerreL */
static double laguerre(int p, int a, double x)
{
   double lm1 = 0.0, l = 1.0;
   for (int k = 0; k < p; ++k)
   {
      const double next = ((2 * k + 1 + a - x) * l - (k + a) * lm1) / (k + 1);
      lm1 = l;
      l = next;
   }
   return l;
}

/* :93 MaskPhase */
static double mask_phase(int doubleHelix, int maskModes, double maskWaist, double rhoNorm, double phi)
{
   if (!doubleHelix)
      return 0.0;
   const int n = imax(2, maskModes);
   const double w = maskWaist > 0.2 ? maskWaist : 0.2;
   const double rc = rhoNorm < 1.0 ? rhoNorm : 1.0;
   const double u = (0.0 < rc ? rc : 0.0) / w;
   const double u2 = u * u;
   double re = 0.0, im = 0.0;
   for (int p = 0; p < n; ++p)
   {
      const int l = 2 * p + 1;
      const double amp = pow(u, (double)l) * exp(-u2) * laguerre(p, l, 2.0 * u2);
      re += amp * cos(l * phi);
      im += amp * sin(l * phi);
   }
   return (re == 0.0 && im == 0.0) ? 0.0 : atan2(im, re);
}

/* :118-130 GeometryFor, :134-148 ZernikePupilSamples */
int32_t orc_psf_geometry(const orc_psf_request* r, int32_t* size, int32_t* nz, int32_t* M)
{
   const int os = imax(1, r->oversampling);
   const int camHalf = imax(1, r->kernelHalfWidthPx);
   const int halfOv = camHalf * os;
   const int sz = 2 * halfOv + 1;
   const int nzWanted = imax(3, r->nz);
   if (size)
      *size = sz;
   if (nz)
      *nz = (nzWanted % 2 == 1) ? nzWanted : nzWanted + 1;
   if (M)
   {
      if (r->pupilSamples > 0)
         *M = imax(8, imin(1024, r->pupilSamples));
      else
      {
         const double diagonalUm = 2.0 * sqrt(2.0) * camHalf * r->pixelSizeNm / 1000.0;
         const double m = 4.0 + 2.0 * r->na * diagonalUm / (r->wavelengthNm / 1000.0);
         const int need = imax(64, imin(512, 4 * (int)ceil(m / 4.0)));
         int fftLen = 1;
         while (fftLen < need + sz - 1)
            fftLen <<= 1;
         *M = imax(need, imin(512, (fftLen - sz + 1) / 4 * 4));
      }
   }
   return 0;
}

/* ComputeZernikePsfPlanes (Sim/ZernikePsf.cpp:150-337) up to the complex field of every plane. Writes either the
 * fields (re, im interleaved, when fields != NULL) or the |E|^2 planes (when planes != NULL). */
static int32_t psf_compute(const orc_psf_request* req, double* fields, float* planes)
{
   if (!(req->na > 0.0) || !(req->wavelengthNm > 0.0) || !(req->immersionIndex > 0.0) || !(req->sampleIndex > 0.0) ||
       !(req->pixelSizeNm > 0.0) || !(req->zStepNm > 0.0))
      return -1;
   int size, nz, kFftM;
   orc_psf_geometry(req, &size, &nz, &kFftM);
   const int os = imax(1, req->oversampling);
   const double resLateralNm = req->pixelSizeNm / os;

   const double NA = req->na;
   const double lambda = req->wavelengthNm * 1E-9;
   const double ni = req->immersionIndex;
   const double ns = req->sampleIndex;
   const double ti0 = req->workingDistanceUm * 1E-6;
   const double particleAxialPosition = req->sampleDepthNm * 1E-9;
   const double resAxialM = req->zStepNm * 1E-9;
   const double resLateralM = resLateralNm * 1E-9;
   const int doubleHelix = req->maskType == 1;

   /* the nonzero terms, in index order (:178-188) */
   int nModes = 0;
   double mc[28];
   int mn[28], ml[28], mm[28];
   for (int j = 0; j < 28; ++j)
   {
      if (req->zernike[j] == 0.0)
         continue;
      mc[nModes] = req->zernike[j];
      index_to_nm(j, &mn[nModes], &ml[nModes]);
      mm[nModes] = abs(ml[nModes]);
      ++nModes;
   }

   const double k0 = 2.0 * kPi / lambda;
   const double bMax = (ns / NA < 1.0) ? ns / NA : 1.0;
   const double kMax = k0 * NA * bMax;
   const double dk = (2.0 * kMax) / (kFftM - 4);
   const double kMin = -floor(kFftM / 2.0) * dk;
   const int nx = size, ny = size;
   const double x0m = -((nx - 1) / 2.0) * resLateralM;
   czt_plan plan;
   czt_init(&plan, kFftM, dk, kMin, nx, resLateralM, x0m);

   const int M = kFftM, c0 = kFftM / 2;
   const size_t MM = (size_t)M * M;
   char* inside = (char*)calloc(MM, 1);
   double* opd1Tab = (double*)calloc(MM, sizeof(double));
   double* opd3Factor = (double*)calloc(MM, sizeof(double));
   double* zernTab = (double*)calloc(MM, sizeof(double));
   double* maskTab = (double*)calloc(MM, sizeof(double));
   for (int iy = 0; iy < M; ++iy)
   {
      const double ky = (iy - c0) * dk;
      for (int ix = 0; ix < M; ++ix)
      {
         const double kx = (ix - c0) * dk;
         const double kr2 = kx * kx + ky * ky;
         if (kr2 > kMax * kMax)
            continue;
         const size_t idx = (size_t)iy * M + ix;
         inside[idx] = 1;
         const double kr = sqrt(kr2);
         const double rho = kr / (k0 * NA);
         const double phi = atan2(ky, kx);
         const double s1 = NA * rho / ns;
         const double s3 = NA * rho / ni;
         const double a1 = 1.0 - s1 * s1, a3 = 1.0 - s3 * s3;
         opd1Tab[idx] = ns * particleAxialPosition * sqrt(0.0 < a1 ? a1 : 0.0);
         opd3Factor[idx] = sqrt(0.0 < a3 ? a3 : 0.0);
         double zernikePhase = 0.0;
         for (int t = 0; t < nModes; ++t)
            zernikePhase += mc[t] * zernike_value(mn[t], ml[t], mm[t], rho / bMax, phi);
         zernikePhase *= 2.0 * kPi;
         zernTab[idx] = zernikePhase;
         maskTab[idx] = mask_phase(doubleHelix, req->maskModes, req->maskWaist, rho / bMax, phi);
      }
   }

   const double focalShift = particleAxialPosition * (ni / ns);
   double* pupilRe = (double*)malloc(sizeof(double) * MM);
   double* pupilIm = (double*)malloc(sizeof(double) * MM);
   double* midRe = (double*)malloc(sizeof(double) * (size_t)M * ny);
   double* midIm = (double*)malloc(sizeof(double) * (size_t)M * ny);
   double* lineRe = (double*)malloc(sizeof(double) * (size_t)M);
   double* lineIm = (double*)malloc(sizeof(double) * (size_t)M);
   double* oRe = (double*)malloc(sizeof(double) * (size_t)imax(nx, ny));
   double* oIm = (double*)malloc(sizeof(double) * (size_t)imax(nx, ny));
   double* aRe = (double*)malloc(sizeof(double) * (size_t)plan.L);
   double* aIm = (double*)malloc(sizeof(double) * (size_t)plan.L);
   for (int z = 0; z < nz; ++z)
   {
      const double ti = (ti0 - focalShift) + resAxialM * (z - (nz - 1.0) / 2.0);
      for (size_t idx = 0; idx < MM; ++idx)
      {
         pupilRe[idx] = 0.0;
         pupilIm[idx] = 0.0;
         if (!inside[idx])
            continue;
         const double opd3 = ni * (ti - ti0) * opd3Factor[idx];
         const double phase = k0 * (opd1Tab[idx] + opd3) + zernTab[idx] + maskTab[idx];
         pupilRe[idx] = cos(phase);
         pupilIm[idx] = sin(phase);
      }
      /* row pass: one CZT over ky per pupil column (:265-312); all-zero columns give exact zeros */
      for (int m = 0; m < M; ++m)
      {
         for (int n = 0; n < M; ++n)
         {
            lineRe[n] = pupilRe[(size_t)n * M + m];
            lineIm[n] = pupilIm[(size_t)n * M + m];
         }
         czt_apply(&plan, lineRe, lineIm, oRe, oIm, aRe, aIm);
         for (int y = 0; y < ny; ++y)
         {
            midRe[(size_t)y * M + m] = oRe[y];
            midIm[(size_t)y * M + m] = oIm[y];
         }
      }
      /* column pass: one CZT over kx per intermediate row (:313-334) */
      for (int y = 0; y < ny; ++y)
      {
         czt_apply(&plan, midRe + (size_t)y * M, midIm + (size_t)y * M, oRe, oIm, aRe, aIm);
         for (int x = 0; x < nx; ++x)
         {
            const size_t o = ((size_t)z * ny + y) * nx + x;
            if (planes)
               planes[o] = (float)(oRe[x] * oRe[x] + oIm[x] * oIm[x]);
            if (fields)
            {
               fields[2 * o] = oRe[x];
               fields[2 * o + 1] = oIm[x];
            }
         }
      }
   }
   free(pupilRe);
   free(pupilIm);
   free(midRe);
   free(midIm);
   free(lineRe);
   free(lineIm);
   free(oRe);
   free(oIm);
   free(aRe);
   free(aIm);
   free(inside);
   free(opd1Tab);
   free(opd3Factor);
   free(zernTab);
   free(maskTab);
   czt_free(&plan);
   return 0;
}

int32_t orc_psf_planes(const orc_psf_request* r, float* out) { return psf_compute(r, NULL, out); }

int32_t orc_psf_fields(const orc_psf_request* r, double* out) { return psf_compute(r, out, NULL); }

/* Sim/ZernikePsf.cpp:401-410 */
void orc_psf_normalize(float* planes, int32_t nz, int32_t n)
{
   const size_t plane = (size_t)n * n;
   for (int32_t z = 0; z < nz; ++z)
   {
      float* p = planes + (size_t)z * plane;
      double sum = 0.0;
      for (size_t i = 0; i < plane; ++i)
         sum += p[i];
      if (sum > 0.0)
         for (size_t i = 0; i < plane; ++i)
            p[i] = (float)(p[i] / sum);
   }
}

/* ======================================================================= */
/* Kernel cache and splat: Sim/PsfGeneratorBridge.cpp, Sim/SplatKernel.inl */
/* ======================================================================= */

/* Sim/PsfGeneratorBridge.cpp:1021-1050 BuildBlockSums */
void orc_block_sums(const float* kernel, int32_t n, int32_t os, float* out)
{
   const int off = os - 1, bw = n + off;
   double* hsum = (double*)malloc(sizeof(double) * (size_t)bw * n);
   for (int y = 0; y < n; ++y)
      for (int bi = 0; bi < bw; ++bi)
      {
         const int b = bi - off, x0 = imax(0, b), x1 = imin(n - 1, b + off);
         double acc = 0.0;
         for (int x = x0; x <= x1; ++x)
            acc += kernel[(size_t)y * n + x];
         hsum[(size_t)y * bw + bi] = acc;
      }
   for (int ai = 0; ai < bw; ++ai)
   {
      const int a = ai - off, y0 = imax(0, a), y1 = imin(n - 1, a + off);
      for (int bi = 0; bi < bw; ++bi)
      {
         double acc = 0.0;
         for (int y = y0; y <= y1; ++y)
            acc += hsum[(size_t)y * bw + bi];
         out[(size_t)ai * bw + bi] = (float)acc;
      }
   }
   free(hsum);
}

/* Sim/PsfGeneratorBridge.cpp:40-60 PsfKernelCache::NearestZIndex */
int32_t orc_nearest_z(int32_t nz, double zStepNm, double zUm, int32_t* clamped)
{
   if (clamped)
      *clamped = 0;
   if (nz <= 1 || zStepNm <= 0.0)
      return 0;
   const double idxF = zUm * 1000.0 / zStepNm + (nz - 1) / 2.0;
   int idx = (int)lround(idxF);
   if (idx < 0)
   {
      idx = 0;
      if (clamped)
         *clamped = 1;
   }
   if (idx >= nz)
   {
      idx = nz - 1;
      if (clamped)
         *clamped = 1;
   }
   return idx;
}

/* :1250 CatmullRomWeights */
static void catmull_rom(double f, double* w)
{
   const double f2 = f * f, f3 = f2 * f;
   w[0] = -0.5 * f3 + f2 - 0.5 * f;
   w[1] = 1.5 * f3 - 2.5 * f2 + 1.0;
   w[2] = -1.5 * f3 + 2.0 * f2 + 0.5 * f;
   w[3] = 0.5 * f3 - 0.5 * f2;
}

/* :1382-1422 SplatSetup (interp 0 nearest, 1 linear, 2 cubic) */
void orc_splat_setup(int32_t sizeOversampled, int32_t os_, double xPx, double yPx, int32_t interp,
                     orc_splat_setup_t* r)
{
   memset(r, 0, sizeof(*r));
   const int os = imax(1, os_);
   const double kc = (sizeOversampled - 1) / 2.0;
   r->x0 = (int)floor(xPx + 0.5);
   r->y0 = (int)floor(yPx + 0.5);
   const double tx = kc + (r->x0 - xPx - 0.5) * os + 0.5;
   const double ty = kc + (r->y0 - yPx - 0.5) * os + 0.5;
   if (interp == 0)
   {
      r->bx = (int)floor(tx + 0.5);
      r->by = (int)floor(ty + 0.5);
      r->nTaps = 1;
      r->wx[0] = 1.0;
      r->wy[0] = 1.0;
      return;
   }
   const int bx = (int)floor(tx), by = (int)floor(ty);
   const double fx = tx - bx, fy = ty - by;
   if (interp == 2)
   {
      r->bx = bx - 1;
      r->by = by - 1;
      r->nTaps = 4;
      catmull_rom(fx, r->wx);
      catmull_rom(fy, r->wy);
      return;
   }
   r->bx = bx;
   r->by = by;
   r->nTaps = 2;
   r->wx[0] = 1.0 - fx;
   r->wx[1] = fx;
   r->wy[0] = 1.0 - fy;
   r->wy[1] = fy;
}

/* Sim/SplatKernel.inl:51-167 (PixelTaps and Rows<NT>: the interior lanes run the same per-pixel sequence), rows
 * [0, height), optional halo spans for this plane. */
static void splat_rows(float* img, uint32_t width, uint32_t height, const float* B, int bw, int os, int camRad,
                       const orc_splat_setup_t* st, double photons, const int16_t* spanLo, const int16_t* spanHi)
{
   const int off = os - 1, NT = st->nTaps;
   const int dyLo = imax(-camRad, -st->y0), dyHi = imin(camRad, (int)height - 1 - st->y0);
   const int dxLo = imax(-camRad, -st->x0), dxHi = imin(camRad, (int)width - 1 - st->x0);
   const int base = st->bx + off;
   for (int dy = dyLo; dy <= dyHi; ++dy)
   {
      const int r0 = st->by + dy * os + off;
      const int jLo = imax(0, -r0), jHi = imin(NT, bw - r0);
      float* rowOut = img + (size_t)(st->y0 + dy) * width;
      int rLo = dxLo, rHi = dxHi;
      if (spanLo)
      {
         rLo = imax(rLo, spanLo[dy + camRad]);
         rHi = imin(rHi, spanHi[dy + camRad]);
         if (rLo > rHi)
            continue;
      }
      for (int dx = rLo; dx <= rHi; ++dx)
      {
         const int c0 = base + dx * os;
         const int iLo = imax(0, -c0), iHi = imin(NT, bw - c0);
         double sum = 0.0;
         for (int j = jLo; j < jHi; ++j)
         {
            double row = 0.0;
            const float* brow = B + (size_t)(r0 + j) * bw;
            for (int i = iLo; i < iHi; ++i)
               row += st->wx[i] * brow[c0 + i];
            sum += st->wy[j] * row;
         }
         rowOut[st->x0 + dx] += (float)(photons * sum);
      }
   }
}

/* Sim/SMLMSimulation.cpp:103-157 RenderPhotonImage, serial path (bands give the same bits, :176-200). */
void orc_render_frame(float* img, uint32_t width, uint32_t height, double bg, const float* bgMap, double bgScale,
                      const orc_emitter* ems, int32_t nEms, const float* blockSums, int32_t nzK, int32_t bw,
                      int32_t os, int32_t sizeOversampled, int32_t interp, const int16_t* spanLo,
                      const int16_t* spanHi, double minPhotons)
{
   const size_t n = (size_t)width * height;
   if (!bgMap && bgScale == 1.0)
   {
      const float v = (float)bg;
      for (size_t i = 0; i < n; ++i)
         img[i] = v;
   }
   else
      for (size_t i = 0; i < n; ++i)
      {
         const double b = bgMap ? (double)bgMap[i] : bg;
         img[i] = (float)(b * bgScale);
      }
   const int camRad = ((sizeOversampled - 1) / 2) / imax(1, os);
   const size_t planeSize = (size_t)bw * bw;
   const int nSpan = 2 * camRad + 1;
   for (int32_t e = 0; e < nEms; ++e)
   {
      const orc_emitter* em = &ems[e];
      if (minPhotons > 0.0 && em->photons < minPhotons)
         continue;
      /* PlanSplat (:1424-1464): no splat for photons <= 0 or a plane outside the stack */
      if (em->photons <= 0.0 || em->zIndex < 0 || em->zIndex >= nzK)
         continue;
      orc_splat_setup_t st;
      orc_splat_setup(sizeOversampled, os, em->xPx, em->yPx, interp, &st);
      splat_rows(img, width, height, blockSums + (size_t)em->zIndex * planeSize, bw, imax(1, os), camRad, &st,
                 em->photons, spanLo ? spanLo + (size_t)em->zIndex * nSpan : NULL,
                 spanHi ? spanHi + (size_t)em->zIndex * nSpan : NULL);
   }
}

/* ---- the Fft placement (PsfInterpMode::Fft) ---------------------------------------------------------------- */

/* Sim/FftRadix2.h:39-69 MakeTwiddles: per stage len the len/2 values of the recurrence from (1, 0); rev = the
 * bit-reversal permutation. c, s: n - 1 doubles; rev: n ints. */
static void radix2_twiddles(int n, int sign, double* c, double* s, int* rev)
{
   for (int len = 2; len <= n; len <<= 1)
   {
      const double ang = sign * 2.0 * 3.14159265358979323846 / len;
      const double wr = cos(ang), wi = sin(ang);
      const int half = len / 2;
      double cr = 1.0, ci = 0.0;
      for (int k = 0; k < half; ++k)
      {
         c[half - 1 + k] = cr;
         s[half - 1 + k] = ci;
         const double nr = cr * wr - ci * wi;
         ci = cr * wi + ci * wr;
         cr = nr;
      }
   }
   rev[0] = 0;
   for (int i = 1, j = 0; i < n; ++i)
   {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1)
         j ^= bit;
      j ^= bit;
      rev[i] = j;
   }
}

/* Sim/FftRadix2.h:74-105 Fft1d: in-place iterative radix-2, unnormalized */
static void radix2_fft(double* re, double* im, int n, const double* c, const double* s, const int* rev)
{
   for (int i = 1; i < n; ++i)
   {
      const int j = rev[i];
      if (i < j)
      {
         double t = re[i];
         re[i] = re[j];
         re[j] = t;
         t = im[i];
         im[i] = im[j];
         im[j] = t;
      }
   }
   for (int len = 2; len <= n; len <<= 1)
   {
      const int half = len / 2;
      const double* tc = c + (half - 1);
      const double* ts = s + (half - 1);
      for (int i = 0; i < n; i += len)
      {
         double* ar = re + i;
         double* ai = im + i;
         double* br = re + i + half;
         double* bi = im + i + half;
         for (int k = 0; k < half; ++k)
         {
            const double cr = tc[k], ci = ts[k];
            const double vr = br[k] * cr - bi[k] * ci, vi = br[k] * ci + bi[k] * cr;
            br[k] = ar[k] - vr;
            bi[k] = ai[k] - vi;
            ar[k] += vr;
            ai[k] += vi;
         }
      }
   }
}

typedef struct
{
   int N;
   double *fc, *fs, *ic, *is;
   int* rev;
} radix2_plan;

/* Sim/PsfGeneratorBridge.cpp:1264-1281 FourierShiftLine */
static void fourier_shift_line(double* re, double* im, const radix2_plan* p, const double* pc, const double* ps)
{
   const int N = p->N;
   radix2_fft(re, im, N, p->fc, p->fs, p->rev);
   for (int k = 0; k < N; ++k)
   {
      const double r = re[k], i = im[k];
      re[k] = r * pc[k] - i * ps[k];
      im[k] = r * ps[k] + i * pc[k];
   }
   radix2_fft(re, im, N, p->ic, p->is, p->rev);
   const double norm = 1.0 / N;
   for (int k = 0; k < N; ++k)
   {
      re[k] *= norm;
      im[k] *= norm;
   }
}

/* Sim/PsfGeneratorBridge.cpp:1293-1380 FftShiftKernelTile: g(i) = f(i + shift) by zero-padded (N = pow2 >= 2n)
 * Fourier shifts along every row, then every column of the complex intermediate; the real part as float. */
void orc_fft_shift_tile(const float* kernel, int32_t n, double shiftX, double shiftY, float* out)
{
   int N = 1;
   while (N < 2 * n)
      N <<= 1;
   const int half = N / 2;
   radix2_plan p;
   p.N = N;
   p.fc = (double*)malloc(sizeof(double) * (size_t)4 * N);
   p.fs = p.fc + N;
   p.ic = p.fs + N;
   p.is = p.ic + N;
   p.rev = (int*)malloc(sizeof(int) * (size_t)N);
   radix2_twiddles(N, -1, p.fc, p.fs, p.rev);
   radix2_twiddles(N, 1, p.ic, p.is, p.rev);
   double* tab = (double*)malloc(sizeof(double) * (size_t)6 * N);
   double *xc = tab, *xs = tab + N, *yc = tab + 2 * N, *ys = tab + 3 * N, *re = tab + 4 * N, *im = tab + 5 * N;
   for (int k = 0; k < N; ++k)
   {
      const int kk = k < half ? k : k - N;
      const double ax = 2.0 * 3.14159265358979323846 * kk * shiftX / N;
      const double ay = 2.0 * 3.14159265358979323846 * kk * shiftY / N;
      xc[k] = cos(ax);
      xs[k] = sin(ax);
      yc[k] = cos(ay);
      ys[k] = sin(ay);
   }
   double* midRe = (double*)malloc(sizeof(double) * (size_t)2 * n * n);
   double* midIm = midRe + (size_t)n * n;
   for (int y = 0; y < n; ++y)
   {
      for (int k = 0; k < N; ++k)
      {
         re[k] = k < n ? kernel[(size_t)y * n + k] : 0.0;
         im[k] = 0.0;
      }
      fourier_shift_line(re, im, &p, xc, xs);
      for (int x = 0; x < n; ++x)
      {
         midRe[(size_t)y * n + x] = re[x];
         midIm[(size_t)y * n + x] = im[x];
      }
   }
   for (int x = 0; x < n; ++x)
   {
      for (int k = 0; k < N; ++k)
      {
         re[k] = k < n ? midRe[(size_t)k * n + x] : 0.0;
         im[k] = k < n ? midIm[(size_t)k * n + x] : 0.0;
      }
      fourier_shift_line(re, im, &p, yc, ys);
      for (int y = 0; y < n; ++y)
         out[(size_t)y * n + x] = (float)re[y];
   }
   free(midRe);
   free(tab);
   free(p.rev);
   free(p.fc);
}

/* RenderPhotonImage with the Fft placement: PlanSplat's Fft branch (Sim/PsfGeneratorBridge.cpp:1443-1456) per
 * emitter -- the nearest setup, ONE Fourier shift of the raw plane by the sub-cell fraction, its block sums read
 * nearest at the rounded grid position. planes: nz raw n x n planes. */
void orc_render_frame_fft(float* img, uint32_t width, uint32_t height, double bg, const float* bgMap, double bgScale,
                          const orc_emitter* ems, int32_t nEms, const float* planes, int32_t nzK, int32_t os_,
                          int32_t n, const int16_t* spanLo, const int16_t* spanHi, double minPhotons)
{
   const int os = imax(1, os_), bw = n + os - 1;
   const size_t npx = (size_t)width * height;
   if (!bgMap && bgScale == 1.0)
   {
      const float v = (float)bg;
      for (size_t i = 0; i < npx; ++i)
         img[i] = v;
   }
   else
      for (size_t i = 0; i < npx; ++i)
      {
         const double b = bgMap ? (double)bgMap[i] : bg;
         img[i] = (float)(b * bgScale);
      }
   const int camRad = ((n - 1) / 2) / os;
   const int nSpan = 2 * camRad + 1;
   float* shifted = (float*)malloc(sizeof(float) * (size_t)n * n);
   float* sums = (float*)malloc(sizeof(float) * (size_t)bw * bw);
   for (int32_t e = 0; e < nEms; ++e)
   {
      const orc_emitter* em = &ems[e];
      if (minPhotons > 0.0 && em->photons < minPhotons)
         continue;
      if (em->photons <= 0.0 || em->zIndex < 0 || em->zIndex >= nzK)
         continue;
      orc_splat_setup_t st;
      orc_splat_setup(n, os, em->xPx, em->yPx, 0, &st);
      const double kc = (n - 1) / 2.0;
      const double tx = kc + (st.x0 - em->xPx - 0.5) * os + 0.5, ty = kc + (st.y0 - em->yPx - 0.5) * os + 0.5;
      const double rx = floor(tx + 0.5), ry = floor(ty + 0.5);
      orc_fft_shift_tile(planes + (size_t)em->zIndex * n * n, n, tx - rx, ty - ry, shifted);
      orc_block_sums(shifted, n, os, sums);
      st.bx = (int)rx;
      st.by = (int)ry;
      splat_rows(img, width, height, sums, bw, os, camRad, &st, em->photons,
                 spanLo ? spanLo + (size_t)em->zIndex * nSpan : NULL, spanHi ? spanHi + (size_t)em->zIndex * nSpan : NULL);
   }
   free(sums);
   free(shifted);
}

/* ======================================================================= */
/* Mean field: Sim/WidefieldRender.cpp:263-323, cli/widefield_check.cpp    */
/* ======================================================================= */

/* :263 ValidUpscale */
static int valid_upscale(int oversampling, int requested)
{
   const int os = imax(1, oversampling);
   for (int u = imax(1, imin(requested, os)); u > 1; --u)
      if (os % u == 0)
         return u;
   return 1;
}

static void wf_geom(int32_t n, int32_t os_, int32_t upscale, int* r, int* s0)
{
   const int os = imax(1, os_);
   *r = imax(1, os / valid_upscale(os, upscale));
   const double kc = (n - 1) / 2.0;
   *s0 = (int)floor(kc - *r / 2.0 + 1.0);
}

/* :297-304 KernelWidefieldPsf::Radius */
int32_t orc_wf_radius(int32_t n, int32_t os, int32_t upscale)
{
   int r, s0;
   wf_geom(n, os, upscale, &r, &s0);
   const int dMin = (int)ceil((-s0 - r + 1) / (double)r);
   const int dMax = (int)floor((n - 1 - s0) / (double)r);
   return imax(1, imax(-dMin, dMax));
}

/* :306-323 KernelWidefieldPsf::Kernel */
void orc_wf_kernel(const float* P, int32_t n, int32_t os, int32_t upscale, int32_t R, float* out)
{
   int r, s0;
   wf_geom(n, os, upscale, &r, &s0);
   const int D = 2 * R + 1;
   for (int dy = -R; dy <= R; ++dy)
   {
      const int ya = s0 + dy * r, yb = imin(n, ya + r);
      for (int dx = -R; dx <= R; ++dx)
      {
         const int xa = s0 + dx * r, xb = imin(n, xa + r);
         double acc = 0.0;
         for (int y = imax(0, ya); y < yb; ++y)
            for (int x = imax(0, xa); x < xb; ++x)
               acc += P[(size_t)y * n + x];
         out[(size_t)(dy + R) * D + (dx + R)] = (float)acc;
      }
   }
}

/* cli/widefield_check.cpp:287-329 DirectImage (before the crop, max and bin), with PlanFocus's plane coordinate
 * (Sim/WidefieldRender.cpp:1476-1490) and KernelWidefieldPsf::PlaneCoord (:285). */
/* PlanFocus's plane blend for one dye plane and focus (Sim/WidefieldRender.cpp:1460-1524): t = the PSF plane
 * coordinate of (centre - focus), clamped to [0, nzP - 1]; plane p0 = floor(t) with weight 1 - fr, plane p0 + 1 with
 * fr (fr = 0 at the last plane). */
void orc_wf_plane_weight(double centreUm, double focusUm, int32_t nzP, double zStepNm, int32_t* p0Out,
                         double* frOut)
{
   const int maxPlane = imax(0, nzP - 1);
   double t = (nzP <= 1) ? 0.0 : (centreUm - focusUm) * 1000.0 / zStepNm + (nzP - 1) / 2.0;
   t = t < 0.0 ? 0.0 : (t > maxPlane ? maxPlane : t);
   int p0 = (int)floor(t);
   double fr = t - p0;
   if (p0 >= maxPlane)
   {
      p0 = maxPlane;
      fr = 0.0;
   }
   *p0Out = p0;
   *frOut = fr;
}

void orc_wf_direct(const float* density, int32_t nx, int32_t ny, int32_t nzD, double dzUm, int64_t k0,
                   const float* kernels, int32_t nzP, int32_t R, double zStepNm, const double* focusUm, int32_t nF,
                   double* out)
{
   const size_t n2 = (size_t)nx * ny;
   const int D = 2 * R + 1;
   memset(out, 0, sizeof(double) * n2 * (size_t)nF);
   for (int32_t f = 0; f < nF; ++f)
   {
      double* img = out + (size_t)f * n2;
      for (int32_t i = 0; i < nzD; ++i)
      {
         const double centre = ((double)(k0 + i) + 0.5) * dzUm;
         int32_t p0;
         double fr;
         orc_wf_plane_weight(centre, focusUm[f], nzP, zStepNm, &p0, &fr);
         for (int s = 0; s < 2; ++s)
         {
            const double w = s ? fr : 1 - fr;
            if (w == 0)
               continue;
            const float* k = kernels + (size_t)(p0 + s) * D * D;
            const float* dens = density + (size_t)i * n2;
            for (int iy = 0; iy < ny; ++iy)
               for (int ix = 0; ix < nx; ++ix)
               {
                  const double a = dens[(size_t)iy * nx + ix];
                  if (a == 0)
                     continue;
                  for (int dy = -R; dy <= R; ++dy)
                  {
                     const int Y = iy + dy;
                     if (Y < 0 || Y >= ny)
                        continue;
                     for (int dx = -R; dx <= R; ++dx)
                     {
                        const int X = ix + dx;
                        if (X < 0 || X >= nx)
                           continue;
                        img[(size_t)Y * nx + X] += w * a * k[(size_t)(dy + R) * D + dx + R];
                     }
                  }
               }
         }
      }
   }
}

/* ======================================================================= */
/* ROI model and its adjoint (FP64)                                         */
/* ======================================================================= */

typedef struct
{
   int M, R, os, Ps, nIn;
   double k0, NA, ni, ns, bMax, kMax, dk, kMin, pix, focalShift, norm;
   double *A, *B, *Z; /* base phase, z slope (per metre), Zernike basis [28][M*M] */
   char* inside;
   double *kx;        /* k_m */
   double *wr, *wi;   /* W[s][m] = exp(i k_m x_s), x_s = (s + 0.5) / os pixels */
} roi_tables;

static void roi_tables_init(roi_tables* t, const orc_roi_model* m, const double* zernike)
{
   t->M = m->pupilSamples;
   t->R = m->roiSize;
   t->os = m->oversampling;
   t->Ps = t->R * t->os;
   const int M = t->M;
   const double lambda = m->wavelengthNm * 1e-9;
   t->NA = m->na;
   t->ni = m->immersionIndex;
   t->ns = m->sampleIndex;
   t->k0 = 2.0 * kPi / lambda;
   t->bMax = (t->ns / t->NA < 1.0) ? t->ns / t->NA : 1.0;
   t->kMax = t->k0 * t->NA * t->bMax;
   t->dk = (2.0 * t->kMax) / (M - 4);
   t->kMin = -floor(M / 2.0) * t->dk;
   t->pix = m->pixelSizeNm * 1e-9;
   const double depth = m->sampleDepthNm * 1e-9;
   t->focalShift = depth * (t->ni / t->ns);
   const size_t MM = (size_t)M * M;
   t->A = (double*)calloc(MM, sizeof(double));
   t->B = (double*)calloc(MM, sizeof(double));
   t->Z = (double*)calloc(28 * MM, sizeof(double));
   t->inside = (char*)calloc(MM, 1);
   t->kx = (double*)malloc(sizeof(double) * (size_t)M);
   const int c0 = M / 2;
   t->nIn = 0;
   for (int j = 0; j < M; ++j)
      t->kx[j] = t->kMin + j * t->dk;
   for (int iy = 0; iy < M; ++iy)
      for (int ix = 0; ix < M; ++ix)
      {
         const double ky = (iy - c0) * t->dk, kx = (ix - c0) * t->dk;
         const double kr2 = kx * kx + ky * ky;
         if (kr2 > t->kMax * t->kMax)
            continue;
         const size_t idx = (size_t)iy * M + ix;
         t->inside[idx] = 1;
         ++t->nIn;
         const double rho = sqrt(kr2) / (t->k0 * t->NA);
         const double phi = atan2(ky, kx);
         const double s1 = t->NA * rho / t->ns, s3 = t->NA * rho / t->ni;
         const double a1 = 1.0 - s1 * s1, a3 = 1.0 - s3 * s3;
         const double opd1 = t->ns * depth * sqrt(a1 > 0.0 ? a1 : 0.0);
         const double opd3 = sqrt(a3 > 0.0 ? a3 : 0.0);
         double zp = 0.0;
         for (int j = 0; j < 28; ++j)
         {
            int n, l;
            index_to_nm(j, &n, &l);
            const double zv = zernike_value(n, l, abs(l), rho / t->bMax, phi);
            t->Z[(size_t)j * MM + idx] = zv;
            zp += zernike[j] * zv;
         }
         t->A[idx] = t->k0 * opd1 + 2.0 * kPi * zp;
         t->B[idx] = t->k0 * t->ni * opd3;
      }
   t->wr = (double*)malloc(sizeof(double) * (size_t)t->Ps * M);
   t->wi = (double*)malloc(sizeof(double) * (size_t)t->Ps * M);
   for (int s = 0; s < t->Ps; ++s)
   {
      const double xs = (s + 0.5) / t->os * t->pix;
      for (int j = 0; j < M; ++j)
      {
         t->wr[(size_t)s * M + j] = cos(t->kx[j] * xs);
         t->wi[(size_t)s * M + j] = sin(t->kx[j] * xs);
      }
   }
   const double q = (t->pix / t->os) * t->dk / (2.0 * kPi);
   t->norm = q * q / t->nIn;
}

static void roi_tables_free(roi_tables* t)
{
   free(t->A);
   free(t->B);
   free(t->Z);
   free(t->inside);
   free(t->kx);
   free(t->wr);
   free(t->wi);
}

/* The shifted pupil P'_mn = P_mn(z) exp(-i (k_m x + k_n y)) of one ROI (re, im: M*M). */
static void roi_pupil(const roi_tables* t, const orc_roi_params* p, double* pr, double* pi)
{
   const int M = t->M;
   const double tz = -t->focalShift + p->z * 1e-9;
   for (int n = 0; n < M; ++n)
      for (int m = 0; m < M; ++m)
      {
         const size_t idx = (size_t)n * M + m;
         if (!t->inside[idx])
         {
            pr[idx] = pi[idx] = 0.0;
            continue;
         }
         const double ph = t->A[idx] + tz * t->B[idx] - (t->kx[m] * p->x + t->kx[n] * p->y) * t->pix;
         pr[idx] = cos(ph);
         pi[idx] = sin(ph);
      }
}

/* E[t][s] = sum_n W[t][n] sum_m P'[n][m] W[s][m] (separable sums). */
static void roi_field(const roi_tables* t, const double* pr, const double* pi, double* er, double* ei, double* tr,
                      double* ti)
{
   const int M = t->M, Ps = t->Ps;
   for (int n = 0; n < M; ++n)
      for (int s = 0; s < Ps; ++s)
      {
         double ar = 0.0, ai = 0.0;
         for (int m = 0; m < M; ++m)
         {
            const double a = pr[(size_t)n * M + m], b = pi[(size_t)n * M + m];
            const double c = t->wr[(size_t)s * M + m], d = t->wi[(size_t)s * M + m];
            ar += a * c - b * d;
            ai += a * d + b * c;
         }
         tr[(size_t)n * Ps + s] = ar;
         ti[(size_t)n * Ps + s] = ai;
      }
   for (int y = 0; y < Ps; ++y)
      for (int s = 0; s < Ps; ++s)
      {
         double ar = 0.0, ai = 0.0;
         for (int n = 0; n < M; ++n)
         {
            const double a = t->wr[(size_t)y * M + n], b = t->wi[(size_t)y * M + n];
            const double c = tr[(size_t)n * Ps + s], d = ti[(size_t)n * Ps + s];
            ar += a * c - b * d;
            ai += a * d + b * c;
         }
         er[(size_t)y * Ps + s] = ar;
         ei[(size_t)y * Ps + s] = ai;
      }
}

void orc_roi_forward(const orc_roi_model* m, const double* zernike, const orc_roi_params* p, int32_t nRoi,
                     double* mu)
{
   roi_tables t;
   roi_tables_init(&t, m, zernike);
   const int M = t.M, Ps = t.Ps, R = t.R, os = t.os;
   double* pr = (double*)malloc(sizeof(double) * (size_t)M * M);
   double* pi = (double*)malloc(sizeof(double) * (size_t)M * M);
   double* tr = (double*)malloc(sizeof(double) * (size_t)M * Ps);
   double* ti = (double*)malloc(sizeof(double) * (size_t)M * Ps);
   double* er = (double*)malloc(sizeof(double) * (size_t)Ps * Ps);
   double* ei = (double*)malloc(sizeof(double) * (size_t)Ps * Ps);
   for (int32_t r = 0; r < nRoi; ++r)
   {
      roi_pupil(&t, p + r, pr, pi);
      roi_field(&t, pr, pi, er, ei, tr, ti);
      double* out = mu + (size_t)r * R * R;
      for (int py = 0; py < R; ++py)
         for (int px = 0; px < R; ++px)
         {
            double acc = 0.0;
            for (int sy = 0; sy < os; ++sy)
               for (int sx = 0; sx < os; ++sx)
               {
                  const size_t o = (size_t)(py * os + sy) * Ps + px * os + sx;
                  acc += er[o] * er[o] + ei[o] * ei[o];
               }
            out[py * R + px] = p[r].photons * t.norm * acc + p[r].background;
         }
   }
   free(pr);
   free(pi);
   free(tr);
   free(ti);
   free(er);
   free(ei);
   roi_tables_free(&t);
}

double orc_roi_backward(const orc_roi_model* m, const double* zernike, const orc_roi_params* p, int32_t nRoi,
                        const double* data, double* gradParams, double* gradZernike)
{
   roi_tables t;
   roi_tables_init(&t, m, zernike);
   const int M = t.M, Ps = t.Ps, R = t.R, os = t.os;
   const size_t MM = (size_t)M * M;
   double* pr = (double*)malloc(sizeof(double) * MM);
   double* pi = (double*)malloc(sizeof(double) * MM);
   double* tr = (double*)malloc(sizeof(double) * (size_t)M * Ps);
   double* ti = (double*)malloc(sizeof(double) * (size_t)M * Ps);
   double* er = (double*)malloc(sizeof(double) * (size_t)Ps * Ps);
   double* ei = (double*)malloc(sizeof(double) * (size_t)Ps * Ps);
   double* w = (double*)malloc(sizeof(double) * (size_t)R * R);
   double* psf = (double*)malloc(sizeof(double) * (size_t)R * R);
   double* br = (double*)malloc(sizeof(double) * MM);
   double* bi = (double*)malloc(sizeof(double) * MM);
   double nll = 0.0;
   for (int j = 0; j < 28; ++j)
      gradZernike[j] = 0.0;
   for (int32_t r = 0; r < nRoi; ++r)
   {
      const orc_roi_params* q = p + r;
      roi_pupil(&t, q, pr, pi);
      roi_field(&t, pr, pi, er, ei, tr, ti);
      const double* d = data + (size_t)r * R * R;
      double gN = 0.0, gb = 0.0;
      for (int py = 0; py < R; ++py)
         for (int px = 0; px < R; ++px)
         {
            double acc = 0.0;
            for (int sy = 0; sy < os; ++sy)
               for (int sx = 0; sx < os; ++sx)
               {
                  const size_t o = (size_t)(py * os + sy) * Ps + px * os + sx;
                  acc += er[o] * er[o] + ei[o] * ei[o];
               }
            const double f = t.norm * acc;
            const double mu = q->photons * f + q->background;
            const double dv = d[py * R + px];
            nll += mu - dv * log(mu);
            const double wv = 1.0 - dv / mu;
            w[py * R + px] = wv;
            psf[py * R + px] = f;
            gN += wv * f;
            gb += wv;
         }
      /* G_s = 2 N norm w_p(s) E_s, into er/ei */
      for (int y = 0; y < Ps; ++y)
         for (int s = 0; s < Ps; ++s)
         {
            const size_t o = (size_t)y * Ps + s;
            const double g = 2.0 * q->photons * t.norm * w[(y / os) * R + s / os];
            er[o] *= g;
            ei[o] *= g;
         }
      /* U[n][s] = sum_y conj(W[y][n]) G[y][s] -> tr/ti */
      for (int n = 0; n < M; ++n)
         for (int s = 0; s < Ps; ++s)
         {
            double ar = 0.0, ai = 0.0;
            for (int y = 0; y < Ps; ++y)
            {
               const double a = t.wr[(size_t)y * M + n], b = -t.wi[(size_t)y * M + n];
               const double c = er[(size_t)y * Ps + s], dd = ei[(size_t)y * Ps + s];
               ar += a * c - b * dd;
               ai += a * dd + b * c;
            }
            tr[(size_t)n * Ps + s] = ar;
            ti[(size_t)n * Ps + s] = ai;
         }
      /* B[n][m] = sum_s U[n][s] conj(W[s][m]) */
      double gx = 0.0, gy = 0.0, gz = 0.0;
      for (int n = 0; n < M; ++n)
         for (int mm = 0; mm < M; ++mm)
         {
            const size_t idx = (size_t)n * M + mm;
            if (!t.inside[idx])
               continue;
            double ar = 0.0, ai = 0.0;
            for (int s = 0; s < Ps; ++s)
            {
               const double a = tr[(size_t)n * Ps + s], b = ti[(size_t)n * Ps + s];
               const double c = t.wr[(size_t)s * M + mm], dd = -t.wi[(size_t)s * M + mm];
               ar += a * c - b * dd;
               ai += a * dd + b * c;
            }
            /* dNLL/dphi = -Im(P' conj(B)) */
            const double g = -(pi[idx] * ar - pr[idx] * ai);
            gx += g * (-t.kx[mm]) * t.pix;
            gy += g * (-t.kx[n]) * t.pix;
            gz += g * t.B[idx] * 1e-9;
            for (int j = 0; j < 28; ++j)
               gradZernike[j] += g * 2.0 * kPi * t.Z[(size_t)j * MM + idx];
         }
      double* gp = gradParams + (size_t)r * 5;
      gp[0] = gx;
      gp[1] = gy;
      gp[2] = gz;
      gp[3] = gN;
      gp[4] = gb;
   }
   free(pr);
   free(pi);
   free(tr);
   free(ti);
   free(er);
   free(ei);
   free(w);
   free(psf);
   free(br);
   free(bi);
   roi_tables_free(&t);
   return nll;
}

/* Sim/WidefieldRender.cpp:547-580 WidefieldImages::Render: per camera pixel the FP64 sum over its u x u cells of
 * max(0, v), v = persistent + sum_j (float)a_j bleach_j in FP32 (j below the last non-zero coefficient), added to
 * cam. channels: [nCh][ch][cw], channel 0 the persistent image when hasPersistent. */
void orc_wf_images_render(const float* channels, int32_t nCh, int32_t hasPersistent, uint32_t cw, uint32_t ch,
                          int32_t upscale, const double* a, int32_t nA, float* cam)
{
   const size_t n = (size_t)cw * ch;
   const float* P = hasPersistent ? channels : NULL;
   const float* Bj = channels + (hasPersistent ? n : 0);
   const int nBleach = nCh - (hasPersistent ? 1 : 0);
   int nb = 0;
   for (int j = 0; j < nBleach && j < nA; ++j)
      if (a[j] != 0.0)
         nb = j + 1;
   if (!P && nb == 0)
      return;
   const unsigned u = upscale > 1 ? (unsigned)upscale : 1u, W = cw / u, H = ch / u;
   float af[64];
   for (int j = 0; j < nb && j < 64; ++j)
      af[j] = (float)a[j];
   for (unsigned Y = 0; Y < H; ++Y)
      for (unsigned X = 0; X < W; ++X)
      {
         double acc = 0.0;
         for (unsigned sy = 0; sy < u; ++sy)
         {
            const size_t row = (size_t)(Y * u + sy) * cw + X * u;
            for (unsigned sx = 0; sx < u; ++sx)
            {
               const size_t i = row + sx;
               float v = P ? P[i] : 0.0f;
               for (int j = 0; j < nb; ++j)
                  v += af[j] * Bj[(size_t)j * n + i];
               acc += v > 0.0f ? v : 0.0f;
            }
         }
         cam[(size_t)Y * W + X] += (float)acc;
      }
}

/* ---- the binned blink regime (Sim/BinnedBlinks.cpp) ------------------------------------------------------------- */

/* BinnedBlinkRenderer::Setup (:94-163): cells per pixel u and kernel radius R. planes null: the in-focus Gaussian
 * (u = clamp(upscale, 1, 8), R = max(1, ceil(3 sigma u))); else KernelWidefieldPsf's u and Radius, cut to
 * ceil((sqrt(haloR2) + 1) u) when haloR2 >= 0 (the halo's largest radius2). */
void orc_binned_geometry(int32_t hasKernel, int32_t size, int32_t os, int32_t upscale, double sigmaPx, int32_t haloR2,
                         int32_t* u, int32_t* R)
{
   if (hasKernel)
   {
      *u = valid_upscale(imax(1, os), upscale);
      int r = orc_wf_radius(size, os, upscale);
      if (haloR2 >= 0)
      {
         const int cut = (int)ceil((sqrt((double)haloR2) + 1.0) * *u);
         r = imin(r, cut);
      }
      *R = imax(1, r);
   }
   else
   {
      *u = imax(1, imin(upscale, 8));
      *R = imax(1, (int)ceil(3.0 * sigmaPx * *u));
   }
}

/* The Gaussian kernel of :146-160: (2R+1)^2 cells exp(-(x^2 + y^2) / (2 (sigma u)^2)), normalised in FP64, then
 * float. */
void orc_binned_gauss(double sigmaPx, int32_t u, int32_t R, float* out)
{
   const int D = 2 * R + 1;
   const double sc = sigmaPx * u, twoS2 = 2.0 * sc * sc;
   double* k = (double*)malloc(sizeof(double) * (size_t)D * D);
   double sum = 0.0;
   for (int y = -R; y <= R; ++y)
      for (int x = -R; x <= R; ++x)
      {
         const double v = exp(-((double)x * x + (double)y * y) / twoS2);
         k[(size_t)(y + R) * D + (x + R)] = v;
         sum += v;
      }
   for (int i = 0; i < D * D; ++i)
      out[i] = (float)(k[i] / sum);
   free(k);
}

/* BinnedBlinkRenderer::Render (:170-246) by the convolution theorem's other side: the per-plane deposits (cell
 * floor((x + 0.5) u) + R, photons summed in float in event order, emitters with photons <= 0 or outside the
 * (W u + 2R) x (H u + 2R) grid skipped, plane clamped to [0, nz - 1]) convolved directly with the plane kernels in
 * FP64 over the cropped FOV cells [R, R + W u) x [R, R + H u); then max(0, (float) cell) summed in FP64 over u x u
 * cells and added (float) to img. kernels: nz x (2R+1)^2 (orc_wf_kernel or orc_binned_gauss). em: n x 4 doubles
 * (xPx, yPx, photons, zIndex). */
void orc_binned_render(const float* kernels, int32_t nz, int32_t u, int32_t R, int32_t W, int32_t H, const double* em,
                       int32_t n, float* img)
{
   const int gw = W * u + 2 * R, gh = H * u + 2 * R, D = 2 * R + 1;
   const int cw = W * u, ch = H * u;
   float* dens = (float*)calloc((size_t)gw * gh, sizeof(float));
   double* cells = (double*)calloc((size_t)cw * ch, sizeof(double));
   int any = 0;
   for (int p = 0; p < nz; ++p)
   {
      int touched = 0;
      for (int i = 0; i < n; ++i)
      {
         const double* e = em + 4 * (size_t)i;
         if (!(e[2] > 0.0))
            continue;
         int pi = (int)e[3];
         pi = imax(0, imin(nz - 1, pi));
         if (pi != p)
            continue;
         const int cx = (int)floor((e[0] + 0.5) * u) + R;
         const int cy = (int)floor((e[1] + 0.5) * u) + R;
         if (cx < 0 || cy < 0 || cx >= gw || cy >= gh)
            continue;
         dens[(size_t)cy * gw + cx] += (float)e[2];
         touched = 1;
      }
      if (!touched)
         continue;
      any = 1;
      const float* K = kernels + (size_t)p * D * D;
      for (int cy = 0; cy < gh; ++cy)
         for (int cx = 0; cx < gw; ++cx)
         {
            const float d = dens[(size_t)cy * gw + cx];
            if (d == 0.0f)
               continue;
            for (int dy = -R; dy <= R; ++dy)
            {
               const int y = cy + dy - R;
               if (y < 0 || y >= ch)
                  continue;
               for (int dx = -R; dx <= R; ++dx)
               {
                  const int x = cx + dx - R;
                  if (x < 0 || x >= cw)
                     continue;
                  cells[(size_t)y * cw + x] += (double)d * K[(size_t)(dy + R) * D + (dx + R)];
               }
            }
            dens[(size_t)cy * gw + cx] = 0.0f;
         }
   }
   if (any)
      for (int y = 0; y < H; ++y)
         for (int x = 0; x < W; ++x)
         {
            double sum = 0.0;
            for (int dy = 0; dy < u; ++dy)
               for (int dx = 0; dx < u; ++dx)
               {
                  const float c = (float)cells[(size_t)(y * u + dy) * cw + (size_t)x * u + dx];
                  sum += c > 0.0f ? c : 0.0f;
               }
            img[(size_t)y * W + x] += (float)sum;
         }
   free(dens);
   free(cells);
}

/* ======================================================================= */
/* JS-compatible FP64 math: the published fdlibm 5.3 algorithms (Sun      */
/* Microsystems, 1993) in the form V8 runs them for Math.* and the         */
/* reference compiles as jsm:: (core/src/jsmath_fdlibm.cpp:137-1564).      */
/* Only the pieces the dye schedules call: log, exp, and sin/cos with the  */
/* argument reductions up to 2^19 pi/2.                                    */
/* ======================================================================= */
/*
 * The algorithms and coefficients restated here are fdlibm's, which carries this notice:
 * ====================================================
 * Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
 *
 * Developed at SunSoft, a Sun Microsystems, Inc. business.
 * Permission to use, copy, modify, and distribute this
 * software is freely granted, provided that this notice
 * is preserved.
 * ====================================================
 * V8's form of them (src/base/ieee754.cc): Copyright 2016 the V8 project authors. All rights reserved.
 */

static int32_t hi_of(double x)
{
   uint64_t u;
   memcpy(&u, &x, 8);
   return (int32_t)(uint32_t)(u >> 32);
}

static uint32_t lo_of(double x)
{
   uint64_t u;
   memcpy(&u, &x, 8);
   return (uint32_t)u;
}

static double of_words(uint32_t hi, uint32_t lo)
{
   const uint64_t u = (uint64_t)hi << 32 | lo;
   double x;
   memcpy(&x, &u, 8);
   return x;
}

/* e_log.c: log(x) = k ln2 + log(1 + f), log(1 + f) = 2s + s R(s^2) with s = f / (2 + f) and R Remez-fitted on
 * [0, 0.1716] (coefficients Lg1..Lg7); k and f from the exponent and mantissa with 1 + f in [sqrt(2)/2, sqrt(2)).
 * V8 form (jsmath_fdlibm.cpp:1278-1357): log(+-0) = -inf, log(<0) = NaN. */
double orc_jsm_log(double x)
{
   const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
   const double two54 = 1.80143985094819840000e+16;
   const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
                Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                Lg7 = 1.479819860511658591e-01;
   int32_t hx = hi_of(x);
   const uint32_t lx = lo_of(x);
   int32_t k = 0;
   if (hx < 0x00100000)
   {
      if (((hx & 0x7fffffff) | (int32_t)lx) == 0)
         return -INFINITY;
      if (hx < 0)
         return NAN;
      k -= 54;
      x *= two54;
      hx = hi_of(x);
   }
   if (hx >= 0x7ff00000)
      return x + x;
   k += (hx >> 20) - 1023;
   hx &= 0x000fffff;
   const int32_t norm = (hx + 0x95f64) & 0x100000;
   x = of_words((uint32_t)(hx | (norm ^ 0x3ff00000)), lo_of(x));
   k += norm >> 20;
   const double f = x - 1.0;
   if ((0x000fffff & (2 + hx)) < 3)
   {
      if (f == 0.0)
      {
         if (k == 0)
            return 0.0;
         const double dk = (double)k;
         return dk * ln2_hi + dk * ln2_lo;
      }
      const double R = f * f * (0.5 - 0.33333333333333333 * f);
      if (k == 0)
         return f - R;
      const double dk = (double)k;
      return dk * ln2_hi - ((R - dk * ln2_lo) - f);
   }
   const double s = f / (2.0 + f);
   const double dk = (double)k;
   const double z = s * s;
   int32_t i = hx - 0x6147a;
   const double w = z * z;
   const int32_t j = 0x6b851 - hx;
   const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
   const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
   i |= j;
   const double R = t2 + t1;
   if (i > 0)
   {
      const double hfsq = 0.5 * f * f;
      if (k == 0)
         return f - (hfsq - s * (hfsq + R));
      return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
   }
   if (k == 0)
      return f - s * (f - R);
   return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

/* e_exp.c: x = k ln2 + r with |r| <= ln2/2 (ln2 split hi + lo), exp(r) = 1 + r + r c / (2 - c) with c = r - r^2
 * P(r^2) (P1..P5 Remez-fitted), scaled by 2^k. V8 form (jsmath_fdlibm.cpp:1135-1276): exp(1) returns the double
 * nearest e, the scale is a multiply by 2^k built from its exponent bits, k = 1024 and k < -1021 split the scale. */
double orc_jsm_exp(double x)
{
   const double halF[2] = {0.5, -0.5};
   const double o_threshold = 7.09782712893383973096e+02, u_threshold = -7.45133219101941108420e+02;
   const double ln2HI[2] = {6.93147180369123816490e-01, -6.93147180369123816490e-01};
   const double ln2LO[2] = {1.90821492927058770002e-10, -1.90821492927058770002e-10};
   const double invln2 = 1.44269504088896338700e+00;
   const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03, P3 = 6.61375632143793436117e-05,
                P4 = -1.65339022054652515390e-06, P5 = 4.13813679705723846039e-08;
   const double E = 2.718281828459045;
   volatile double huge = 1.0e+300, twom1000 = 9.33263618503218878990e-302, two1023 = 8.988465674311579539e307;
   double hi = 0.0, lo = 0.0;
   int32_t k = 0;
   uint32_t hx = (uint32_t)hi_of(x);
   const int32_t xsb = (int32_t)((hx >> 31) & 1);
   hx &= 0x7fffffff;
   if (hx >= 0x40862e42)
   {
      if (hx >= 0x7ff00000)
      {
         if (((hx & 0xfffff) | lo_of(x)) != 0)
            return x + x;
         return xsb == 0 ? x : 0.0;
      }
      if (x > o_threshold)
         return huge * huge;
      if (x < u_threshold)
         return twom1000 * twom1000;
   }
   if (hx > 0x3fd62e42)
   {
      if (hx < 0x3ff0a2b2)
      {
         if (x == 1.0)
            return E;
         hi = x - ln2HI[xsb];
         lo = ln2LO[xsb];
         k = 1 - xsb - xsb;
      }
      else
      {
         k = (int32_t)(invln2 * x + halF[xsb]);
         const double t = k;
         hi = x - t * ln2HI[0];
         lo = t * ln2LO[0];
      }
      x = hi - lo;
   }
   else if (hx < 0x3e300000)
   {
      if (huge + x > 1.0)
         return 1.0 + x;
   }
   else
      k = 0;
   const double t = x * x;
   double twopk;
   if (k >= -1021)
      twopk = of_words(0x3ff00000u + ((uint32_t)k << 20), 0);
   else
      twopk = of_words(0x3ff00000u + ((uint32_t)(k + 1000) << 20), 0);
   const double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
   if (k == 0)
      return 1.0 - ((x * c) / (c - 2.0) - x);
   const double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
   if (k >= -1021)
   {
      if (k == 1024)
         return y * 2.0 * two1023;
      return y * twopk;
   }
   return y * twopk * twom1000;
}

/* High words of n pi/2, n = 1..32 (e_rem_pio2.c's npio2_hw): a reduction of such an x cancels, so it takes the
 * second step. */
static const int32_t kNpio2Hw[32] = {
   0x3ff921fb, 0x400921fb, 0x4012d97c, 0x401921fb, 0x401f6a7a, 0x4022d97c, 0x4025fdbb, 0x402921fb,
   0x402c463a, 0x402f6a7a, 0x4031475c, 0x4032d97c, 0x40346b9c, 0x4035fdbb, 0x40378fdb, 0x403921fb,
   0x403ab41b, 0x403c463a, 0x403dd85a, 0x403f6a7a, 0x40407e4c, 0x4041475c, 0x4042106c, 0x4042d97c,
   0x4043a28c, 0x40446b9c, 0x404534ac, 0x4045fdbb, 0x4046c6cb, 0x40478fdb, 0x404858eb, 0x404921fb};

/* e_rem_pio2.c for |x| <= 2^19 pi/2: x = n pi/2 + (y0 + y1), pi/2 in 33-bit pieces (pio2_1, pio2_2, pio2_3 and the
 * tails) applied one, two or three at a time as cancellation demands. Larger |x|: NaN (the Payne-Hanek branch is
 * not restated). */
static int32_t rem_pio2(double x, double* y)
{
   const double invpio2 = 6.36619772367581382433e-01;
   const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
   const double pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21;
   const double pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32;
   const int32_t hx = hi_of(x);
   const int32_t ix = hx & 0x7fffffff;
   if (ix <= 0x3fe921fb)
   {
      y[0] = x;
      y[1] = 0.0;
      return 0;
   }
   if (ix < 0x4002d97c)
   {
      const double s = hx > 0 ? 1.0 : -1.0;
      double z = x - s * pio2_1;
      if (ix != 0x3ff921fb)
      {
         y[0] = z - s * pio2_1t;
         y[1] = (z - y[0]) - s * pio2_1t;
      }
      else
      {
         z -= s * pio2_2;
         y[0] = z - s * pio2_2t;
         y[1] = (z - y[0]) - s * pio2_2t;
      }
      return hx > 0 ? 1 : -1;
   }
   if (ix > 0x413921fb)
   {
      y[0] = y[1] = NAN;
      return 0;
   }
   const double t0 = fabs(x);
   const int32_t n = (int32_t)(t0 * invpio2 + 0.5);
   const double fn = (double)n;
   double r = t0 - fn * pio2_1;
   double w = fn * pio2_1t;
   y[0] = r - w;
   if (!(n < 32 && ix != kNpio2Hw[n - 1]))
   {
      const int32_t j = ix >> 20;
      int32_t i = j - ((hi_of(y[0]) >> 20) & 0x7ff);
      if (i > 16)
      {
         double t = r;
         w = fn * pio2_2;
         r = t - w;
         w = fn * pio2_2t - ((t - r) - w);
         y[0] = r - w;
         i = j - ((hi_of(y[0]) >> 20) & 0x7ff);
         if (i > 49)
         {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y[0] = r - w;
         }
      }
   }
   y[1] = (r - y[0]) - w;
   if (hx < 0)
   {
      y[0] = -y[0];
      y[1] = -y[1];
      return -n;
   }
   return n;
}

/* k_cos.c: cos(x + y) on [-pi/4, pi/4], 1 - x^2/2 + x^4 C(x^2) (C1..C6), with the x^2/2 term split off against
 * qx (x/4 from the high word, or 0.28125 above 0.78125) for |x| >= 0.3. */
static double kcos(double x, double y)
{
   const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03, C3 = 2.48015872894767294178e-05,
                C4 = -2.75573143513906633035e-07, C5 = 2.08757232129817482790e-09,
                C6 = -1.13596475577881948265e-11;
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix < 0x3e400000 && (int)x == 0)
      return 1.0;
   const double z = x * x;
   const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
   if (ix < 0x3fd33333)
      return 1.0 - (0.5 * z - (z * r - x * y));
   const double qx = ix > 0x3fe90000 ? 0.28125 : of_words((uint32_t)(ix - 0x00200000), 0);
   const double hz = 0.5 * z - qx;
   const double a = 1.0 - qx;
   return a - (hz - (z * r - x * y));
}

/* k_sin.c: sin(x + y) on [-pi/4, pi/4], x + x^3 (S1 + x^2 S(x^2)) (S1..S6); iy = 0 when y is 0. */
static double ksin(double x, double y, int iy)
{
   const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03, S3 = -1.98412698298579493134e-04,
                S4 = 2.75573137070700676789e-06, S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix < 0x3e400000 && (int)x == 0)
      return x;
   const double z = x * x;
   const double v = z * x;
   const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
   if (iy == 0)
      return x + v * (S1 + z * r);
   return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

/* s_cos.c / s_sin.c: reduce by n pi/2, then the kernel of quadrant n mod 4. */
double orc_jsm_cos(double x)
{
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix <= 0x3fe921fb)
      return kcos(x, 0.0);
   if (ix >= 0x7ff00000)
      return x - x;
   double y[2];
   const int32_t n = rem_pio2(x, y);
   switch (n & 3)
   {
   case 0: return kcos(y[0], y[1]);
   case 1: return -ksin(y[0], y[1], 1);
   case 2: return -kcos(y[0], y[1]);
   default: return ksin(y[0], y[1], 1);
   }
}

double orc_jsm_sin(double x)
{
   const int32_t ix = hi_of(x) & 0x7fffffff;
   if (ix <= 0x3fe921fb)
      return ksin(x, 0.0, 0);
   if (ix >= 0x7ff00000)
      return x - x;
   double y[2];
   const int32_t n = rem_pio2(x, y);
   switch (n & 3)
   {
   case 0: return ksin(y[0], y[1], 1);
   case 1: return kcos(y[0], y[1]);
   case 2: return -ksin(y[0], y[1], 1);
   default: return -kcos(y[0], y[1]);
   }
}

/* ======================================================================= */
/* Dye schedules: core/src/dyes.cpp:254-321 (LogNormal, Poisson,           */
/* DyeSchedule), :440-478 (LabelSchedule), :537-657 (PersistentGen and its */
/* two callers), for one set of kinetics (no history). The channel        */
/* numbers are core/src/dyes.h:45-57.                                      */
/* ======================================================================= */

enum
{
   CH_ACT = 4,
   CH_PERSIST = 5,
   CH_INIT_ON = 7,
   CH_AUX = 8,
   CH_SCHED0 = 16,
   CH_SCHED_STRIDE = 8,
   MAX_BLINKS = 1000
};

static double unit_of(uint32_t a) { return ((double)(a >> 9) + 0.5) * 1.1920928955078125e-7; }

static double dye_u(uint32_t h1, int32_t k, int32_t n, uint32_t ch)
{
   uint32_t v[4] = {h1, (uint32_t)k, (uint32_t)n, ch};
   orc_pcg4d(v);
   return unit_of(v[0]);
}

static const double kJsPi = 3.141592653589793;

/* LogNormalFor / LogNormalMean1 (dyes.cpp:257-268): mean 1, CV cv, Box-Muller normal from (u1, u2). */
static void lognormal_for(double cv, double* sigma, double* mu)
{
   const double s2 = orc_jsm_log(1 + cv * cv);
   *sigma = sqrt(s2);
   *mu = -s2 / 2;
}

static double lognormal_mean1(double sigma, double mu, double u1, double u2)
{
   const double z = sqrt(-2 * orc_jsm_log(u1)) * orc_jsm_cos(2 * kJsPi * u2);
   return orc_jsm_exp(mu + sigma * z);
}

/* PoissonInverse / PoissonFromUniform (dyes.cpp:270-290). */
static long poisson_inverse(double m, double expM, double u)
{
   double p = expM, F = p;
   long c = 0;
   while (u > F && c < 1000)
   {
      c++;
      p *= m / c;
      F += p;
   }
   return c;
}

static long poisson_from_uniform(double m, double u, double u2)
{
   if (!(m > 0))
      return 0;
   if (m > 30)
   {
      const double z = sqrt(-2 * orc_jsm_log(u)) * orc_jsm_cos(2 * kJsPi * u2);
      const long c = (long)floor(m + sqrt(m) * z + 0.5);
      return c > 0 ? c : 0;
   }
   return poisson_inverse(m, orc_jsm_exp(-m), u);
}

static void put_blink(double* out, int32_t cap, int32_t i, double tOn, double tOff, double b)
{
   if (i < cap)
   {
      out[3 * i] = tOn;
      out[3 * i + 1] = tOff;
      out[3 * i + 2] = b;
   }
}

/* DyeScheduleFrom (dyes.cpp:300-321): blinks from time t, appended at out[count..]; returns the new count. */
static int32_t schedule_from(uint32_t h1, int32_t k, int32_t n, const orc_kinetics* kin, double t, double tMax,
                             double* out, int32_t cap, int32_t count)
{
   const double pBleach = fmin(1.0, fmax(0.01, kin->bleachProb));
   const double cv = fmax(0.0, kin->photonCV);
   double sigma, mu;
   lognormal_for(cv, &sigma, &mu);
   for (int j = 0; j < MAX_BLINKS; j++)
   {
      if (t >= tMax)
         break;
      const uint32_t base = CH_SCHED0 + (uint32_t)j * CH_SCHED_STRIDE;
      const double on = -orc_jsm_log(dye_u(h1, k, n, base + 0)) * kin->onSec;
      double b = 1;
      if (cv > 0)
      {
         const double u1 = dye_u(h1, k, n, base + 1);
         const double u2 = dye_u(h1, k, n, base + 2);
         b = lognormal_mean1(sigma, mu, u1, u2);
      }
      put_blink(out, cap, count++, t, t + on, b);
      if (dye_u(h1, k, n, base + 3) < pBleach)
         break;
      t += on - orc_jsm_log(dye_u(h1, k, n, base + 4)) * kin->offSec;
   }
   return count;
}

/* LabelSchedule's blinks (dyes.cpp:440-471): dSTORM shifts DyeSchedule by the initial ON draw, PALM runs it from
 * 0, DNA-PAINT and WideField have none here. */
int32_t orc_label_blinks(uint32_t h1, int32_t k, int32_t n, int32_t mode, const orc_kinetics* kin, double tMax,
                         double* out, int32_t cap)
{
   if (mode == 2 || mode == 3)
      return 0;
   double shift = 0.0;
   if (mode == 0 && kin->initialOnSec > 0)
      shift = -orc_jsm_log(dye_u(h1, k, n, CH_INIT_ON)) * kin->initialOnSec;
   /* DyeSchedule (dyes.cpp:293-298) with tMax - shift */
   if (!(kin->activationRatePerSec > 0))
      return 0;
   const double t = -orc_jsm_log(dye_u(h1, k, n, CH_ACT)) / kin->activationRatePerSec;
   const int32_t count = schedule_from(h1, k, n, kin, t, tMax - shift, out, cap, 0);
   if (shift != 0)
      for (int32_t i = 0; i < count && i < cap; ++i)
      {
         out[3 * i] += shift;
         out[3 * i + 1] += shift;
      }
   return count;
}

/* LabelSchedule's continuous windows (dyes.cpp:447-477): WideField ALWAYS_ON (3) with the AUX draw, dSTORM
 * INITIAL_ON (2) for a nonzero initial ON, PALM with a pre state PRE (1) until the first activation. */
int32_t orc_label_cont(uint32_t h1, int32_t k, int32_t n, int32_t mode, int32_t preState, const orc_kinetics* kin,
                       double w[4])
{
   if (mode == 3)
   {
      w[0] = 0.0;
      w[1] = INFINITY;
      w[2] = 3;
      w[3] = -orc_jsm_log(dye_u(h1, k, n, CH_AUX));
      return 1;
   }
   if (mode == 0 && kin->initialOnSec > 0)
   {
      const double shift = -orc_jsm_log(dye_u(h1, k, n, CH_INIT_ON)) * kin->initialOnSec;
      if (shift != 0)
      {
         w[0] = 0.0;
         w[1] = shift;
         w[2] = 2;
         w[3] = 0.0;
         return 1;
      }
      return 0;
   }
   if (mode == 1 && preState)
   {
      w[0] = 0.0;
      w[1] = kin->activationRatePerSec > 0
                ? -orc_jsm_log(dye_u(h1, k, n, CH_ACT)) / kin->activationRatePerSec
                : INFINITY;
      w[2] = 1;
      w[3] = -orc_jsm_log(dye_u(h1, k, n, CH_AUX));
      return 1;
   }
   return 0;
}

/* PersistentGen (dyes.cpp:566-626) for one set of kinetics, with PersistentBlinks' window filter and bin range
 * (dyes.cpp:636-646) or PersistentBlinksInBins' unfiltered bins (dyes.cpp:648-655). */
int32_t orc_persistent_blinks(uint32_t h1, int32_t k, int32_t n, const orc_kinetics* kin, int32_t bins, double t0,
                              double t1, int64_t binLo, int64_t binHi, double* out, int32_t cap)
{
   const double maxOn = 20.0 * kin->onSec;
   long b0 = (long)binLo, b1 = (long)binHi;
   if (!bins)
   {
      if (!(t1 > t0))
         return 0;
      const long f = (long)floor((t0 - maxOn) / 1.0);
      b0 = f > 0 ? f : 0;
      b1 = (long)floor(t1 / 1.0);
   }
   if (!(kin->activationRatePerSec > 0))
      return 0;
   uint32_t v[4] = {h1, (uint32_t)k, (uint32_t)n, CH_PERSIST};
   orc_pcg4d(v);
   const uint32_t key = v[0];
   /* BinKineticsOf (dyes.cpp:544-558) */
   const double rate = kin->activationRatePerSec, onSec = kin->onSec;
   const double cv = fmax(0.0, kin->photonCV);
   double sigma, mu;
   lognormal_for(cv, &sigma, &mu);
   const double m = rate * 1.0;
   const int small = m <= 30;
   const double expM = small ? orc_jsm_exp(-m) : 0.0;
   int32_t count = 0;
   for (long b = b0 > 0 ? b0 : 0; b <= b1; b++)
   {
      const uint32_t bin = (uint32_t)b;
      const long c = small ? poisson_inverse(m, expM, dye_u(key, (int32_t)bin, 0, 0))
                           : poisson_from_uniform(m, dye_u(key, (int32_t)bin, 0, 0), dye_u(key, (int32_t)bin, 0, 1));
      for (long j = 0; j < c; j++)
      {
         const uint32_t jj = (uint32_t)j;
         const double u = dye_u(key, (int32_t)bin, (int32_t)jj, 2);
         const double tOn = (bin + u) * 1.0;
         double on = -orc_jsm_log(dye_u(key, (int32_t)bin, (int32_t)jj, 3)) * onSec;
         on = fmin(maxOn, on);
         if (!bins && !(tOn < t1 && tOn + on > t0))
            continue;
         double br = 1;
         if (cv > 0)
         {
            const double u1 = dye_u(key, (int32_t)bin, (int32_t)jj, 4);
            const double u2 = dye_u(key, (int32_t)bin, (int32_t)jj, 5);
            br = lognormal_mean1(sigma, mu, u1, u2);
         }
         if (count < cap)
         {
            double* o = out + 5 * (size_t)count;
            o[0] = tOn;
            o[1] = tOn + on;
            o[2] = br;
            o[3] = bin;
            o[4] = jj;
         }
         count++;
      }
   }
   return count;
}
