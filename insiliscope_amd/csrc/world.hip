// world.hip -- the cell field's dye blocks and window queries on the device (isc::World's block walk, core/src/
// world.cpp:455-640, with the dye generation and the queries it drives), bit-identical to the reference's.
//
// The caller's cell geometry stays on the device (HBM holds every cell of a large field several times over). At
// creation: the microtubules' parallel-transport frames (BuildMtFrames) and reach, one lane per microtubule; the
// candidate dye blocks in the walk's order (per cell: every 1 um block of every microtubule, then the pores' and the
// lamina's 1 um cubes, EnvelopeBlocks) with their midpoints. A query then runs, on the context's stream:
//   select   one lane per candidate block: the walk's reach tests against the window (cell, microtubule, block);
//   dyes     one workgroup per selected block: its sites' label and fluorescence draws, counted, then (after a scan)
//            the fluorescent ones' positions (LinkedDye on the frames, LocalToWorld, the 0.1 nm lattice of
//            World::DyeStore) written in the reference's site order;
//   query    sites: the dyes in the window, compacted in order; density: one atomic add per dye (integer counts,
//            exact in any order); continuous: one window per dye at most; events: each dye's blinks (the schedule
//            state machines of dye_gen.h, one lane per dye), then each dSTORM / PALM block's blinks sorted by
//            (tOn, dye, blink) -- bitonic in LDS, merge passes above 4096 -- and gathered into event rows.
// Every step is integer work or FP64 in the reference's operation order (the file is compiled without FP
// contraction; jsm.h restates the transcendentals), so positions and times are the reference's bits.
#include "dye_gen.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <vector>

using namespace iscg;
using namespace iscg::dyegen;

namespace {

// core/src/dyes.h:24-41, core/src/npcs.h:28-34
constexpr double kNm = 1e-3;
constexpr double kMtRadiusNm = 12.5;
constexpr int kProto = 13;
constexpr double kDimerNm = 8;
constexpr double kLatticeStart = 3;
constexpr uint32_t kSeamPhaseCh = 8000000u;
constexpr uint32_t kDyeSalt = 0x9E3779B9u, kNpcDyeSalt = 0x4E504344u, kLaminDyeSalt = 0x4C414D44u;
constexpr double kDyeBlockUm = 1.0, kNpcBlockUm = 1.0;
constexpr int kNpcSites = 64;
constexpr uint32_t CH_LABEL = 0, CH_LINK_U = 1, CH_LINK_PHI = 2, CH_LINK_R = 3, CH_FLUOR = 6;
constexpr uint32_t CH_BIND_U = 13, CH_BIND_PHI = 14, CH_BIND_R = 15;
// World::DyeStore (core/src/world.h:394-405): positions on a 0.1 nm lattice offset by kPhase of a step
constexpr double kStepsPerUm = 10000.0;
constexpr double kPhase = 0.3819660112501051;
constexpr int kStructures = 4;  // microtubules, pores, lamina, docking strands (STRUCTURE_COUNT)

// Candidate dimers per protofilament in a 1 um block: the walk starts one dimer before the block (floor - 1), and a
// block holds at most 1000 / 8 = 125 dimers of a protofilament, so 128 slots cover every site the loop can take.
constexpr int kMtSlots = 128;
constexpr int kMtCands = kProto * kMtSlots;

constexpr int kThreads = 256;     // dye-block workgroups, selection, scans
constexpr int kSortThreads = 512;
constexpr int kSortMax = 4096;    // one LDS sort per segment up to this many blinks

struct Pt
{
   double x, y, z;
};

// One microtubule: its points are [p0, p0 + nPts) of the point arrays (frames too: cum at each point, T, U, V per
// segment at its first point).
struct MtDev
{
   int64_t p0;
   int32_t nPts, cell, index, pad;
   double len, reach;
};

// A candidate dye block: kind 0 a microtubule's block `block` (elem: the microtubule), 1 pores / 2 lamina (elem: the
// envelope block); its midpoint (world x, y; z) for the walk's tests.
struct Cand
{
   int32_t cell, kind, elem, block;
   double wx, wy, z;
};

// An envelope block: the pores (or lamina points) ids[first .. first + count) of a cell whose centres lie in one
// cell-local cube (x, y, z: its centre).
struct EnvBlock
{
   int32_t cell, first, count, pad;
   double x, y, z;
};

// A selected block as the dye kernels and queries read it.
struct BlockMeta
{
   uint32_t h1;
   int32_t structure, cell, mtIndex;
};

// A dye as the queries read it: its lattice position, id and address kn = k << 24 | n (World::PackedDye).
struct DyeRec
{
   double x, y, z;
   uint32_t id, kn;
};

// One structure's label as the kernels use it (normalised linkage, LinkageSetup's cubes, the clamped fractions).
struct LabelDev
{
   double below, ff;  // density and fluorescent fraction clamped to [0, 1]
   int32_t mode, preState;
   iscg_kinetics kin;
   double binderNm, flexNm, flexUm, minL3, maxL3;
};

struct Window
{
   double x0, y0, x1, y1, zMin, zMax;
};

// What a query's selection and dye kernels read.
struct QueryArgs
{
   Window win;
   uint32_t seed;
   int32_t npcOn, laminOn;
   double mtReach, blockReach, npcRb, laminRb;
   double npcSiteReach;
   LabelDev lab[3];
   HistDev hist[3];  // each structure's kinetics history of two or more segments (n = 0: the label's kinetics)
};

// ---- device helpers ----------------------------------------------------------------------------------------------

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }  // std::min

// RectDist (core/src/world.cpp:38-43): distance from (px, py) to the window (0 inside).
__device__ __forceinline__ double rect_dist(double px, double py, const Window& w)
{
   const double dx = dmax(dmax(w.x0 - px, 0.0), px - w.x1);
   const double dy = dmax(dmax(w.y0 - py, 0.0), py - w.y1);
   return jsm::hypot(dx, dy);
}

// LocalToWorld (core/src/world.cpp:169-176)
__device__ __forceinline__ void local_to_world(const iscg_world_cell& c, double lx, double ly, double& wx,
                                               double& wy)
{
   const double rot = c.packRot;
   if (rot == 0 || isnan(rot))
   {
      wx = c.x + lx;
      wy = c.y + ly;
      return;
   }
   const double cr = jsm::cos(rot), sr = jsm::sin(rot);
   wx = c.x + lx * cr - ly * sr;
   wy = c.y + lx * sr + ly * cr;
}

// MtSegmentAt (core/src/dyes.cpp:17-27): the first segment whose end reaches arc S, clamped to the last.
__device__ __forceinline__ int seg_at(const double* __restrict__ cum, int nSeg, double S)
{
   int lo = 0, hi = nSeg - 1;
   while (lo < hi)
   {
      const int mid = (lo + hi) / 2;
      if (cum[mid + 1] < S)
         lo = mid + 1;
      else
         hi = mid;
   }
   return lo;
}

// World::DyeStore: the lattice value of a coordinate (Steps, then Um).
__device__ __forceinline__ double on_lattice(double um)
{
   const long long g = (long long)floor(um * kStepsPerUm - kPhase + 0.5);
   return ((double)g + kPhase) / kStepsPerUm;
}

// Pcg4d(a, b, c, d).a
__device__ __forceinline__ uint32_t hash_a(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
   pcg4d(a, b, c, d);
   return a;
}

// LinkedDye (core/src/dyes.cpp:93-118): the dye of a site with base b, unit axis a and the site's draws.
__device__ Pt linked_dye(const Pt& b, const Pt& a, double attachNm, const LabelDev& L, const Address& at, double r1,
                         double r2, double r3)
{
   const double B = (attachNm + L.binderNm) * kNm;
   double tx = b.x + a.x * B, ty = b.y + a.y * B, tz = b.z + a.z * B;
   if (L.flexNm > 0)
   {
      const double fu = at.draw(CH_BIND_U) * 2 - 1;
      const double fphi = at.draw(CH_BIND_PHI) * 2 * jsm::kPi;
      const double fr = L.flexUm * jsm::cbrt(at.draw(CH_BIND_R));
      const double fs = sqrt(1 - fu * fu);
      tx += fr * fs * jsm::cos(fphi);
      ty += fr * fs * jsm::sin(fphi);
      tz += fr * fu;
   }
   const double lu = r1 * 2 - 1, phi = r2 * 2 * jsm::kPi, sn = sqrt(1 - lu * lu);
   const double r = jsm::cbrt(L.minL3 + (L.maxL3 - L.minL3) * r3);
   return Pt{tx + r * sn * jsm::cos(phi), ty + r * sn * jsm::sin(phi), tz + r * lu};
}

// ---- creation kernels --------------------------------------------------------------------------------------------

// BuildMtFrames (core/src/microtubules.cpp:701-725) and the microtubule's reach (EnsureMts, core/src/world.cpp:
// 80-95: the farthest point from the cell centre in xy), one lane per microtubule.
__global__ __launch_bounds__(kThreads) void frames_kernel(const Pt* __restrict__ pts, MtDev* __restrict__ mts,
                                                          int32_t nMts, double* __restrict__ cum, Pt* __restrict__ T,
                                                          Pt* __restrict__ U, Pt* __restrict__ V)
{
   const int m = blockIdx.x * kThreads + threadIdx.x;
   if (m >= nMts)
      return;
   MtDev& mt = mts[m];
   const Pt* p = pts + mt.p0;
   double* c = cum + mt.p0;
   const int n = mt.nPts;
   double reach = 0;
   for (int i = 0; i < n; ++i)
      reach = dmax(reach, jsm::hypot(p[i].x, p[i].y));
   mt.reach = reach;
   if (n < 2)
   {
      if (n == 1)
         c[0] = 0;
      mt.len = 0;
      return;
   }
   Pt* t_ = T + mt.p0;
   Pt* u_ = U + mt.p0;
   Pt* v_ = V + mt.p0;
   c[0] = 0;
   Pt prevU{0, 0, 0}, prevT{1, 0, 0};
   for (int i = 0; i + 1 < n; ++i)
   {
      const double dx = p[i + 1].x - p[i].x, dy = p[i + 1].y - p[i].y, dz = p[i + 1].z - p[i].z;
      const double len = jsm::hypot(dx, dy, dz);
      c[i + 1] = c[i] + len;
      Pt t;
      if (len > 1e-12)
         t = Pt{dx / len, dy / len, dz / len};
      else
         t = i > 0 ? prevT : Pt{1, 0, 0};
      Pt u = i == 0 ? (fabs(t.z) < 0.9 ? Pt{0, 0, 1} : Pt{1, 0, 0}) : prevU;
      const double d = u.x * t.x + u.y * t.y + u.z * t.z;
      u = Pt{u.x - d * t.x, u.y - d * t.y, u.z - d * t.z};
      double ul = jsm::hypot(u.x, u.y, u.z);
      if (ul == 0 || isnan(ul))  // NormOr1
         ul = 1;
      u = Pt{u.x / ul, u.y / ul, u.z / ul};
      t_[i] = t;
      u_[i] = u;
      v_[i] = Pt{t.y * u.z - t.z * u.y, t.z * u.x - t.x * u.z, t.x * u.y - t.y * u.x};
      prevU = u;
      prevT = t;
   }
   mt.len = c[n - 1];
}

// The candidate blocks' midpoints: a microtubule block's arc midpoint (CellAssets::Mids, core/src/world.cpp:183-204:
// PointAtArc at min((b + 0.5) um, length)), an envelope block's cube centre, each in world x, y.
__global__ __launch_bounds__(kThreads) void mids_kernel(Cand* __restrict__ cands, int32_t n,
                                                        const iscg_world_cell* __restrict__ cells,
                                                        const MtDev* __restrict__ mts, const Pt* __restrict__ pts,
                                                        const double* __restrict__ cum, const Pt* __restrict__ T,
                                                        const EnvBlock* __restrict__ npcB,
                                                        const EnvBlock* __restrict__ lamB)
{
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n)
      return;
   Cand& c = cands[i];
   const iscg_world_cell cell = cells[c.cell];
   double lx, ly, lz;
   if (c.kind == 0)
   {
      const MtDev mt = mts[c.elem];
      const double S = dmin((c.block + 0.5) * kDyeBlockUm, mt.len);
      const double* cm = cum + mt.p0;
      const int seg = seg_at(cm, mt.nPts - 1, S);
      const double f = S - cm[seg];
      const Pt p = pts[mt.p0 + seg], t = T[mt.p0 + seg];
      lx = p.x + t.x * f;
      ly = p.y + t.y * f;
      lz = p.z + t.z * f;
   }
   else
   {
      const EnvBlock b = c.kind == 1 ? npcB[c.elem] : lamB[c.elem];
      lx = b.x;
      ly = b.y;
      lz = b.z;
   }
   local_to_world(cell, lx, ly, c.wx, c.wy);
   c.z = lz;
}

// The farthest pore or lamina point centre from the cell centre in xy (EnvelopeBlocks' reachXY), one workgroup per
// cell (a maximum: the same in any order).
__global__ __launch_bounds__(kThreads) void env_reach_kernel(const int32_t* __restrict__ start, const Pt* __restrict__ c0,
                                                             int strideDoubles, int32_t nCells, double* __restrict__ out)
{
   __shared__ double part[kThreads / 64];
   const int cidx = blockIdx.x;
   double r = 0;
   for (int i = start[cidx] + threadIdx.x; i < start[cidx + 1]; i += kThreads)
   {
      const double* q = reinterpret_cast<const double*>(c0) + (size_t)i * strideDoubles;
      r = dmax(r, jsm::hypot(q[0], q[1]));
   }
   for (int d = 32; d > 0; d >>= 1)
      r = dmax(r, __shfl_xor(r, d, 64));
   if ((threadIdx.x & 63) == 0)
      part[threadIdx.x >> 6] = r;
   __syncthreads();
   if (threadIdx.x == 0)
   {
      double m = 0;
      for (int w = 0; w < kThreads / 64; ++w)
         m = dmax(m, part[w]);
      out[cidx] = m;
   }
}

// ---- scans -------------------------------------------------------------------------------------------------------

constexpr int kScanPer = 8;                       // counts per lane
constexpr int kScanTile = kThreads * kScanPer;    // counts per workgroup

__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const int32_t* __restrict__ v, int64_t n,
                                                               long long* __restrict__ part)
{
   const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
   long long s = 0;
#pragma unroll
   for (int r = 0; r < kScanPer; ++r)
      s += b + r < n ? v[b + r] : 0;
   long long tot;
   (void)block_exclusive<kThreads / 64>(s, tot);
   if (threadIdx.x == 0)
      part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_top_kernel(long long* __restrict__ part, int32_t nPart,
                                                        int64_t* __restrict__ off, int64_t n)
{
   long long carry = 0;
   for (int base = 0; base < nPart; base += 1024)
   {
      const int i = base + threadIdx.x;
      const long long v = i < nPart ? part[i] : 0;
      long long chunk;
      const long long ex = block_exclusive<16>(v, chunk);
      if (i < nPart)
         part[i] = carry + ex;
      carry += chunk;
   }
   if (threadIdx.x == 0)
      off[n] = carry;
}

__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const int32_t* __restrict__ v, int64_t n,
                                                              const long long* __restrict__ part,
                                                              int64_t* __restrict__ off)
{
   const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
   int32_t c[kScanPer];
   long long s = 0;
#pragma unroll
   for (int r = 0; r < kScanPer; ++r)
   {
      c[r] = b + r < n ? v[b + r] : 0;
      s += c[r];
   }
   long long tot;
   long long o = part[blockIdx.x] + block_exclusive<kThreads / 64>(s, tot);
#pragma unroll
   for (int r = 0; r < kScanPer; ++r)
   {
      if (b + r < n)
         off[b + r] = o;
      o += c[r];
   }
}

// ---- query kernels -----------------------------------------------------------------------------------------------

// The walk's tests (World::ForEachDyeBlock, core/src/world.cpp:556-589) for each candidate block: 1 if it is walked.
__global__ __launch_bounds__(kThreads) void select_kernel(const Cand* __restrict__ cands, int32_t n,
                                                          const iscg_world_cell* __restrict__ cells,
                                                          const MtDev* __restrict__ mts,
                                                          const double* __restrict__ npcReach,
                                                          const double* __restrict__ laminReach, QueryArgs q,
                                                          int32_t* __restrict__ flag)
{
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n)
      return;
   const Cand c = cands[i];
   const iscg_world_cell cell = cells[c.cell];
   const Window& w = q.win;
   int32_t on = 0;
   // CellsInRectPtr: the cell's footprint circle meets the window
   const double centreDist = rect_dist(cell.x, cell.y, w);
   if (centreDist <= cell.rOuter)
   {
      double rb;
      bool reach;
      if (c.kind == 0)
      {
         rb = q.blockReach;
         reach = !(centreDist > mts[c.elem].reach + q.mtReach);
      }
      else if (c.kind == 1)
      {
         rb = q.npcRb;
         reach = q.npcOn && !(centreDist > npcReach[c.cell] + rb);
      }
      else
      {
         rb = q.laminRb;
         reach = q.laminOn && !(centreDist > laminReach[c.cell] + rb);
      }
      on = reach && !(c.z + rb < w.zMin || c.z - rb >= w.zMax) && !(rect_dist(c.wx, c.wy, w) > rb);
   }
   flag[i] = on;
}

__global__ __launch_bounds__(kThreads) void compact_kernel(const int32_t* __restrict__ flag,
                                                           const int64_t* __restrict__ off, int64_t n,
                                                           int32_t* __restrict__ out)
{
   const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < n && flag[i])
      out[off[i]] = (int32_t)i;
}

// What the dye kernels need of a selected block.
struct BlockView
{
   int kind, nCand, block;
   uint32_t h1;
   MtDev mt;
   EnvBlock env;
   double blockNm0;
};

struct Geo
{
   const iscg_world_cell* cells;
   const Cand* cands;
   const MtDev* mts;
   const Pt* pts;
   const double* cum;
   const Pt* T;
   const Pt* U;
   const Pt* V;
   const EnvBlock* npcB;
   const EnvBlock* lamB;
   const int32_t* npcIds;
   const int32_t* lamIds;
   const iscg_npc* npcs;
   const iscg_lamin_point* lamin;
   const double* npcSites;
   const int32_t* npcStart;
   const int32_t* laminStart;
   double off[kProto];  // MtProtofilamentOffsetNm
};

__device__ __forceinline__ BlockView block_view(const Geo& g, const QueryArgs& q, int32_t candIdx)
{
   const Cand c = g.cands[candIdx];
   const iscg_world_cell cell = g.cells[c.cell];
   BlockView v{};
   v.kind = c.kind;
   v.block = c.block;
   if (c.kind == 0)
   {
      v.mt = g.mts[c.elem];
      v.h1 = hash_a(q.seed ^ kDyeSalt, (uint32_t)cell.cx, (uint32_t)cell.cy, (uint32_t)v.mt.index);
      v.blockNm0 = c.block * kDyeBlockUm * 1000;
      const LabelDev& L = q.lab[0];
      // DyesInBlock's early returns
      const bool none = v.mt.nPts < 2 || c.block < 0 || v.blockNm0 * kNm >= v.mt.len || !(L.below > 0) || !(L.ff > 0);
      v.nCand = none ? 0 : kMtCands;
   }
   else
   {
      v.env = c.kind == 1 ? g.npcB[c.elem] : g.lamB[c.elem];
      const uint32_t salt = c.kind == 1 ? kNpcDyeSalt : kLaminDyeSalt;
      v.h1 = hash_a(q.seed ^ salt, (uint32_t)cell.cx, (uint32_t)cell.cy, 0u);
      const LabelDev& L = q.lab[c.kind];
      const bool none = !(L.below > 0) || !(L.ff > 0);
      v.nCand = none ? 0 : v.env.count * (c.kind == 1 ? kNpcSites : 1);
   }
   return v;
}

// Site j of a block: its address (k, n) and whether the walk takes it and its dye is fluorescent.
struct Site
{
   int32_t k, n, elem;  // elem: the pore or lamina point (global index)
   double S;            // arc length (microtubule)
   bool on;
};

__device__ __forceinline__ Site site_of(const Geo& g, const QueryArgs& q, const BlockView& v, int j)
{
   Site s{0, 0, 0, 0.0, false};
   const LabelDev& L = q.lab[v.kind];
   if (v.kind == 0)
   {
      const int k = j / kMtSlots, t = j % kMtSlots;
      const double off = g.off[k];
      const long long n0 = (long long)floor((v.blockNm0 - off) / kDimerNm) - 1;
      const long long n = (n0 > 0 ? n0 : 0) + t;
      const double sNm = off + n * kDimerNm;
      const double blk = floor(sNm / (kDyeBlockUm * 1000));
      if (blk != (double)v.block)
         return s;
      const double S = sNm * kNm;
      if (S >= v.mt.len)
         return s;
      s.k = k;
      s.n = (int32_t)n;
      s.S = S;
   }
   else if (v.kind == 1)
   {
      const int32_t p = g.npcIds[v.env.first + j / kNpcSites];
      s.elem = p;
      s.k = j % kNpcSites;
      s.n = g.npcs[p].cand;
   }
   else
   {
      const int32_t p = g.lamIds[v.env.first + j];
      s.elem = p;
      s.k = 0;
      s.n = g.lamin[p].cand;
   }
   const Address at = Address::of(v.h1, (uint32_t)s.k, (uint32_t)s.n);
   if (at.draw(CH_LABEL) >= L.below)
      return s;
   if (L.ff < 1.0 && at.draw(CH_FLUOR) >= L.ff)
      return s;
   s.on = true;
   return s;
}

__global__ __launch_bounds__(kThreads) void block_count_kernel(Geo g, QueryArgs q, const int32_t* __restrict__ sel,
                                                               int32_t nSel, int32_t* __restrict__ counts,
                                                               BlockMeta* __restrict__ meta)
{
   const int s = blockIdx.x;
   const BlockView v = block_view(g, q, sel[s]);
   long long mine = 0;
   for (int j = threadIdx.x; j < v.nCand; j += kThreads)
      mine += site_of(g, q, v, j).on;
   long long tot;
   (void)block_exclusive<kThreads / 64>(mine, tot);
   if (threadIdx.x == 0)
   {
      counts[s] = (int32_t)tot;
      const Cand c = g.cands[sel[s]];
      meta[s] = BlockMeta{v.h1, c.kind, c.cell, c.kind == 0 ? v.mt.index : -1};
   }
}

// What a density query bins into (Density3dInWindow's grid and scales).
struct DensityArgs
{
   float* out;
   unsigned long long* total;
   int32_t nx, ny, nz;
   uint32_t mask;
   double sx, sy, sz;
};

constexpr int kFillRecords = 0, kFillRaw = 1, kFillDensity = 2;
// Voxels of a block's local counts in LDS (a box around its midpoint holding every dye it can make).
constexpr int kLocalBins = 4096;

// MODE kFillRecords: the dyes as the world keeps them (world coordinates on the lattice) with their block;
// kFillRaw: the generators' own output (cell-local, unrounded) as iscg_block_dye rows, those below capacity;
// kFillDensity: no records, each dye in the window counted into its voxel (Density3dInWindow), through LDS counts
// over a box around the block's midpoint that every dye of the block falls in (its reach), added to the grid once
// per voxel (integer counts: the same in any order).
template <int MODE>
__global__ __launch_bounds__(kThreads) void block_fill_kernel(Geo g, QueryArgs q, const int32_t* __restrict__ sel,
                                                              const int64_t* __restrict__ dyeStart,
                                                              DyeRec* __restrict__ dyes, int32_t* __restrict__ dyeBlock,
                                                              iscg_block_dye* __restrict__ raw, int64_t capacity,
                                                              DensityArgs den)
{
   const int s = blockIdx.x;
   const BlockView v = block_view(g, q, sel[s]);
   const Cand cand = g.cands[sel[s]];
   const int32_t cellIdx = cand.cell;
   const iscg_world_cell cell = g.cells[cellIdx];
   // the density box: voxels [b0, b0 + bn) per axis (x, y, z), used when it fits kLocalBins
   __shared__ int box[7];
   __shared__ float bins[MODE == kFillDensity ? kLocalBins : 1];
   long long counted = 0;
   bool binThis = false;
   if constexpr (MODE == kFillDensity)
   {
      binThis = ((den.mask >> v.kind) & 1u) != 0;
      if (threadIdx.x == 0)
      {
         const double rad = v.kind == 0 ? q.blockReach : v.kind == 1 ? q.npcRb : q.laminRb;
         const Window& w = q.win;
         auto range = [&](double c, double lo, double sc, int n, int& b0, int& bn) {
            const double f0 = fmax(-2.0, fmin((double)n + 1, floor((c - rad - lo) * sc) - 1));
            const double f1 = fmax(-2.0, fmin((double)n + 1, floor((c + rad - lo) * sc) + 1));
            const int i0 = max(0, (int)f0), i1 = min(n - 1, (int)f1);
            b0 = i0;
            bn = i1 >= i0 ? i1 - i0 + 1 : 0;
         };
         range(cand.wx, w.x0, den.sx, den.nx, box[0], box[3]);
         range(cand.wy, w.y0, den.sy, den.ny, box[1], box[4]);
         if (den.nz > 1)
            range(cand.z, w.zMin, den.sz, den.nz, box[2], box[5]);
         else
         {
            box[2] = 0;
            box[5] = 1;
         }
         const long long nb = (long long)box[3] * box[4] * box[5];
         box[6] = nb <= kLocalBins ? (int)nb : -1;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < max(box[6], 0); i += kThreads)
         bins[i] = 0.0f;
   }
   __shared__ double sc[16], ss[16];  // per protofilament (cos, sin theta), or per pore corner
   __shared__ double rotc, rots;
   __shared__ int rotated;
   if (threadIdx.x == 0)
   {
      const double rot = cell.packRot;
      rotated = !(rot == 0 || isnan(rot));
      rotc = rotated ? jsm::cos(rot) : 1.0;
      rots = rotated ? jsm::sin(rot) : 0.0;
   }
   if (v.kind == 0 && threadIdx.x < kProto)
   {
      // MtSeamPhase and MtProtofilamentTheta (core/src/dyes.cpp:29-39)
      const uint32_t a = hash_a(q.seed, (uint32_t)cell.cx, (uint32_t)cell.cy, kSeamPhaseCh + (uint32_t)v.mt.index);
      const double phase = 2 * jsm::kPi * CounterRng::unit_of(a);
      const double theta = phase + (double)((int)threadIdx.x * 2) * jsm::kPi / kProto;
      sc[threadIdx.x] = jsm::cos(theta);
      ss[threadIdx.x] = jsm::sin(theta);
   }
   else if (v.kind == 1 && threadIdx.x < 8)
   {
      const double ang = (double)(int)threadIdx.x * jsm::kPi / 4;
      sc[threadIdx.x] = jsm::cos(ang);
      ss[threadIdx.x] = jsm::sin(ang);
   }
   __syncthreads();
   const LabelDev& L = q.lab[v.kind];
   long long base = MODE == kFillDensity ? 0 : dyeStart[s];
   for (int j0 = 0; j0 < v.nCand; j0 += kThreads)
   {
      const int j = j0 + threadIdx.x;
      Site st{};
      if (j < v.nCand)
         st = site_of(g, q, v, j);
      long long tot = 0, ex = 0;
      if constexpr (MODE != kFillDensity)
         ex = block_exclusive<kThreads / 64>(st.on ? 1 : 0, tot);
      if (st.on && (MODE != kFillDensity || binThis))
      {
         const Address at = Address::of(v.h1, (uint32_t)st.k, (uint32_t)st.n);
         const double r1 = at.draw(CH_LINK_U), r2 = at.draw(CH_LINK_PHI), r3 = at.draw(CH_LINK_R);
         Pt pos;
         if (v.kind == 0)
         {
            const int nSeg = v.mt.nPts - 1;
            const double* cm = g.cum + v.mt.p0;
            const int seg = seg_at(cm, nSeg, st.S);
            const Pt ft = g.T[v.mt.p0 + seg], fu = g.U[v.mt.p0 + seg], fv = g.V[v.mt.p0 + seg];
            const Pt p = g.pts[v.mt.p0 + seg];
            const double f = st.S - cm[seg];
            const Pt c{p.x + ft.x * f, p.y + ft.y * f, p.z + ft.z * f};
            const double ct = sc[st.k], stt = ss[st.k];
            const Pt a{ct * fu.x + stt * fv.x, ct * fu.y + stt * fv.y, ct * fu.z + stt * fv.z};
            pos = linked_dye(c, a, kMtRadiusNm, L, at, r1, r2, r3);
         }
         else if (v.kind == 1)
         {
            const iscg_npc& pq = g.npcs[st.elem];
            const double sa = g.npcSites[3 * st.k] * kNm, sb = g.npcSites[3 * st.k + 1] * kNm,
                         sc3 = g.npcSites[3 * st.k + 2] * kNm;
            const Pt atp{pq.c[0] + sa * pq.e1[0] + sb * pq.e2[0] + sc3 * pq.n[0],
                         pq.c[1] + sa * pq.e1[1] + sb * pq.e2[1] + sc3 * pq.n[1],
                         pq.c[2] + sa * pq.e1[2] + sb * pq.e2[2] + sc3 * pq.n[2]};
            const int corner = (st.k & 31) >> 2;
            const double ct = sc[corner], stt = ss[corner];
            const Pt a{ct * pq.e1[0] + stt * pq.e2[0], ct * pq.e1[1] + stt * pq.e2[1], ct * pq.e1[2] + stt * pq.e2[2]};
            pos = linked_dye(atp, a, 0, L, at, r1, r2, r3);
         }
         else
         {
            const iscg_lamin_point& pq = g.lamin[st.elem];
            // the binder's direction, drawn at the unused site address k = 1 (core/src/dyes.cpp:241-245)
            const Address a1 = Address::of(v.h1, 1u, (uint32_t)st.n);
            const double au = a1.draw(CH_LINK_U) * 2 - 1;
            const double aphi = a1.draw(CH_LINK_PHI) * 2 * jsm::kPi;
            const double as = sqrt(1 - au * au);
            const Pt a{as * jsm::cos(aphi), as * jsm::sin(aphi), au};
            pos = linked_dye(Pt{pq.c[0], pq.c[1], pq.c[2]}, a, 0, L, at, r1, r2, r3);
         }
         const long long d = base + ex;
         if constexpr (MODE == kFillRaw)
         {
            if (d < capacity)
            {
               const int32_t elem = v.kind == 0   ? v.mt.index
                                    : v.kind == 1 ? st.elem - g.npcStart[cellIdx]
                                                  : st.elem - g.laminStart[cellIdx];
               raw[d] = iscg_block_dye{pos.x, pos.y, pos.z, at.word(CH_LABEL), s, v.kind, elem, st.k, st.n, 0};
            }
         }
         else
         {
         // GenerateDyes (core/src/world.cpp:808-838): LocalToWorld with the block's cos/sin, then the lattice
         double wx, wy;
         if (!rotated)
         {
            wx = cell.x + pos.x;
            wy = cell.y + pos.y;
         }
         else
         {
            wx = cell.x + pos.x * rotc - pos.y * rots;
            wy = cell.y + pos.x * rots + pos.y * rotc;
         }
         const DyeRec rec{on_lattice(wx), on_lattice(wy), on_lattice(pos.z), at.word(CH_LABEL),
                          ((uint32_t)st.k << 24) | ((uint32_t)st.n & 0xFFFFFFu)};
         if constexpr (MODE == kFillRecords)
         {
            dyes[d] = rec;
            dyeBlock[d] = s;
         }
         else
         {
            const Window& w = q.win;
            if (rec.z >= w.zMin && rec.z < w.zMax && rec.x >= w.x0 && rec.x < w.x1 && rec.y >= w.y0 && rec.y < w.y1)
            {
               const int ix = min(den.nx - 1, (int)floor((rec.x - w.x0) * den.sx));
               const int iy = min(den.ny - 1, (int)floor((rec.y - w.y0) * den.sy));
               const int iz = den.nz > 1 ? min(den.nz - 1, (int)floor((rec.z - w.zMin) * den.sz)) : 0;
               const int lx = ix - box[0], ly = iy - box[1], lz = iz - box[2];
               if (box[6] > 0 && lx >= 0 && lx < box[3] && ly >= 0 && ly < box[4] && lz >= 0 && lz < box[5])
                  atomicAdd(bins + (lz * box[4] + ly) * box[3] + lx, 1.0f);
               else
                  atomicAdd(den.out + ((size_t)iz * den.ny + iy) * den.nx + ix, 1.0f);
               ++counted;
            }
         }
         }
      }
      base += tot;
   }
   if constexpr (MODE == kFillDensity)
   {
      __syncthreads();
      const int nb = box[6];
      for (int i = threadIdx.x; i < nb; i += kThreads)
      {
         const float c = bins[i];
         if (c != 0.0f)
         {
            const int lx = i % box[3], ly = (i / box[3]) % box[4], lz = i / (box[3] * box[4]);
            atomicAdd(den.out + ((size_t)(box[2] + lz) * den.ny + (box[1] + ly)) * den.nx + (box[0] + lx), c);
         }
      }
      long long tot;
      (void)block_exclusive<kThreads / 64>(counted, tot);
      if (threadIdx.x == 0 && tot)
         atomicAdd(den.total, (unsigned long long)tot);
   }
}

__device__ __forceinline__ bool in_window(const DyeRec& d, const Window& w)
{
   return d.z >= w.zMin && d.z < w.zMax && d.x >= w.x0 && d.x < w.x1 && d.y >= w.y0 && d.y < w.y1;
}

// 1 for each dye in the window (of a structure in mask).
__global__ __launch_bounds__(kThreads) void win_flag_kernel(const DyeRec* __restrict__ dyes,
                                                            const int32_t* __restrict__ dyeBlock,
                                                            const BlockMeta* __restrict__ meta, int64_t n, Window w,
                                                            uint32_t mask, int32_t* __restrict__ flag)
{
   const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i >= n)
      return;
   const int s = meta[dyeBlock[i]].structure;
   flag[i] = ((mask >> s) & 1u) && in_window(dyes[i], w);
}

__global__ __launch_bounds__(kThreads) void sites_kernel(const DyeRec* __restrict__ dyes,
                                                         const int32_t* __restrict__ dyeBlock,
                                                         const BlockMeta* __restrict__ meta,
                                                         const iscg_world_cell* __restrict__ cells,
                                                         const int32_t* __restrict__ flag,
                                                         const int64_t* __restrict__ off, int64_t n, int64_t capacity,
                                                         iscg_world_dye* __restrict__ out)
{
   const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i >= n || !flag[i] || off[i] >= capacity)
      return;
   const DyeRec d = dyes[i];
   const BlockMeta m = meta[dyeBlock[i]];
   const iscg_world_cell c = cells[m.cell];
   out[off[i]] = iscg_world_dye{d.x,   d.y, d.z, d.id, c.cx, c.cy, m.mtIndex, (int32_t)(d.kn >> 24),
                                (int32_t)(d.kn & 0xFFFFFFu), m.structure, 0};
}

__global__ __launch_bounds__(kThreads) void zero_kernel(float* __restrict__ out, int64_t n,
                                                        unsigned long long* __restrict__ total)
{
   for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
      out[i] = 0.0f;
   if (blockIdx.x == 0 && threadIdx.x == 0)
      *total = 0;
}

// The in-window dyes' addresses and structures, compacted in dye order (winIdx: their dye index).
__global__ __launch_bounds__(kThreads) void win_list_kernel(const DyeRec* __restrict__ dyes,
                                                            const int32_t* __restrict__ dyeBlock,
                                                            const BlockMeta* __restrict__ meta,
                                                            const int32_t* __restrict__ flag,
                                                            const int64_t* __restrict__ off, int64_t n,
                                                            iscg_dye_address* __restrict__ addr,
                                                            uint8_t* __restrict__ structure,
                                                            int32_t* __restrict__ winIdx)
{
   const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i >= n || !flag[i])
      return;
   const int64_t o = off[i];
   const DyeRec d = dyes[i];
   const BlockMeta m = meta[dyeBlock[i]];
   addr[o] = iscg_dye_address{m.h1, (int32_t)(d.kn >> 24), (int32_t)(d.kn & 0xFFFFFFu)};
   structure[o] = (uint8_t)m.structure;
   winIdx[o] = (int32_t)i;
}

__global__ __launch_bounds__(kThreads) void zero_counts_kernel(int32_t* __restrict__ v, int64_t n)
{
   for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
      v[i] = 0;
}

struct OfStructure
{
   const uint8_t* s;
   uint8_t want;
   __device__ __forceinline__ bool operator()(int i) const { return s[i] == want; }
};

template <class Gen>
__global__ __launch_bounds__(kDyeThreads) void world_count_kernel(Gen g, const iscg_dye_address* __restrict__ addr,
                                                                  int32_t n, int32_t* __restrict__ counts,
                                                                  OfStructure sel)
{
   const int begin = blockIdx.x * kWgDyes + (threadIdx.x >> 6) * kChunk;
   const int end = min(begin + kChunk, n);
   (void)wave_schedules<false>(g, addr, begin, end, counts, nullptr, 0, nullptr, 0, sel);
}

template <class Gen>
__global__ __launch_bounds__(kDyeThreads) void world_fill_kernel(Gen g, const iscg_dye_address* __restrict__ addr,
                                                                 int32_t n, int32_t* __restrict__ counts,
                                                                 const int64_t* __restrict__ off,
                                                                 typename Gen::Record* __restrict__ out,
                                                                 int64_t capacity, OfStructure sel)
{
   const int begin = blockIdx.x * kWgDyes + (threadIdx.x >> 6) * kChunk;
   const int end = min(begin + kChunk, n);
   (void)wave_schedules<true>(g, addr, begin, end, counts, reinterpret_cast<const long long*>(off), 0, out,
                              capacity, sel);
}

// A sort key: (tOn, the blink's index in its block): the order of a stable sort by tOn.
__device__ __forceinline__ bool key_less(double ta, int32_t ia, double tb, int32_t ib)
{
   return ta < tb || (ta == tb && ia < ib);
}

// The event range of selected block s: its in-window dyes [winOff[dyeStart[s]], winOff[dyeStart[s + 1]]) and their
// blinks.
__device__ __forceinline__ void seg_of(int s, const int64_t* dyeStart, const int64_t* winOff, const int64_t* evOff,
                                       int64_t& e0, int64_t& e1)
{
   e0 = evOff[winOff[dyeStart[s]]];
   e1 = evOff[winOff[dyeStart[s + 1]]];
}

// Per selected block: the permutation of its blinks into event order. Blocks of a DNA-PAINT label keep the
// schedule order (dye, bin, index); the others are sorted by (tOn, dye, blink), in LDS in chunks of kSortMax (the
// merge passes join the chunks of longer blocks).
__global__ __launch_bounds__(kSortThreads) void seg_sort_kernel(const BlockMeta* __restrict__ meta, QueryArgs q,
                                                                const int64_t* __restrict__ dyeStart,
                                                                const int64_t* __restrict__ winOff,
                                                                const int64_t* __restrict__ evOff,
                                                                const iscg_blink* __restrict__ blinks,
                                                                int32_t* __restrict__ perm)
{
   __shared__ double kt[kSortMax];
   __shared__ int32_t ki[kSortMax];
   const int s = blockIdx.x;
   int64_t e0, e1;
   seg_of(s, dyeStart, winOff, evOff, e0, e1);
   const int32_t len = (int32_t)(e1 - e0);
   const int st = meta[s].structure;
   const bool sorted = st < 3 && (q.lab[st].mode == 0 || q.lab[st].mode == 1);
   if (!sorted || len <= 1)
   {
      for (int32_t i = threadIdx.x; i < len; i += kSortThreads)
         perm[e0 + i] = i;
      return;
   }
   for (int32_t c0 = 0; c0 < len; c0 += kSortMax)
   {
      const int32_t m = min(kSortMax, len - c0);
      int32_t p2 = 2;
      while (p2 < m)
         p2 <<= 1;
      for (int32_t i = threadIdx.x; i < p2; i += kSortThreads)
      {
         kt[i] = i < m ? blinks[e0 + c0 + i].tOn : __builtin_inf();
         ki[i] = i < m ? c0 + i : 0x7fffffff;
      }
      __syncthreads();
      for (int32_t size = 2; size <= p2; size <<= 1)
      {
         for (int32_t stride = size >> 1; stride > 0; stride >>= 1)
         {
            for (int32_t t = threadIdx.x; t < p2 / 2; t += kSortThreads)
            {
               const int32_t lo = 2 * t - (t & (stride - 1));
               const int32_t hi = lo + stride;
               const bool up = (lo & size) == 0;
               const double a = kt[lo], b = kt[hi];
               const int32_t ia = ki[lo], ib = ki[hi];
               if (key_less(b, ib, a, ia) == up)
               {
                  kt[lo] = b;
                  kt[hi] = a;
                  ki[lo] = ib;
                  ki[hi] = ia;
               }
            }
            __syncthreads();
         }
      }
      for (int32_t i = threadIdx.x; i < m; i += kSortThreads)
         perm[e0 + c0 + i] = ki[i];
      __syncthreads();
   }
}

// One merge pass over the sorted runs of length R within every sorted block longer than R: each blink's place in
// the merged run is its place in its own run plus the number of keys of the partner run below it.
__global__ __launch_bounds__(kThreads) void merge_pass_kernel(const int32_t* __restrict__ evBlock, int64_t nEv,
                                                              const BlockMeta* __restrict__ meta, QueryArgs q,
                                                              const int64_t* __restrict__ dyeStart,
                                                              const int64_t* __restrict__ winOff,
                                                              const int64_t* __restrict__ evOff,
                                                              const iscg_blink* __restrict__ blinks, int32_t R,
                                                              const int32_t* __restrict__ in, int32_t* __restrict__ out)
{
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= nEv)
      return;
   const int s = evBlock[e];
   int64_t e0, e1;
   seg_of(s, dyeStart, winOff, evOff, e0, e1);
   const int32_t len = (int32_t)(e1 - e0), p = (int32_t)(e - e0);
   const int st = meta[s].structure;
   const bool sorted = st < 3 && (q.lab[st].mode == 0 || q.lab[st].mode == 1);
   if (!sorted || len <= R)
   {
      out[e] = in[e];
      return;
   }
   const int32_t run = p / R, runStart = run * R;
   const int32_t me = in[e];
   const double tm = blinks[e0 + me].tOn;
   int32_t pLo, pHi, dest0;
   if ((run & 1) == 0)
   {
      pLo = runStart + R;
      pHi = min(runStart + 2 * R, len);
      dest0 = runStart;
   }
   else
   {
      pLo = runStart - R;
      pHi = runStart;
      dest0 = runStart - R;
   }
   // keys of [pLo, pHi) below (tm, me)
   int32_t lo = pLo, hi = pHi > pLo ? pHi : pLo;
   while (lo < hi)
   {
      const int32_t mid = (lo + hi) >> 1;
      const int32_t o = in[e0 + mid];
      if (key_less(blinks[e0 + o].tOn, o, tm, me))
         lo = mid + 1;
      else
         hi = mid;
   }
   out[e0 + dest0 + (p - runStart) + (lo - pLo)] = me;
}

// The block of each blink (the in-window dye's block), for the merge passes.
__global__ __launch_bounds__(kThreads) void ev_block_kernel(const int64_t* __restrict__ evOff, int32_t nWin,
                                                            const int32_t* __restrict__ winIdx,
                                                            const int32_t* __restrict__ dyeBlock,
                                                            int32_t* __restrict__ evBlock)
{
   const int w = blockIdx.x * kThreads + threadIdx.x;
   if (w >= nWin)
      return;
   const int32_t s = dyeBlock[winIdx[w]];
   for (int64_t e = evOff[w]; e < evOff[w + 1]; ++e)
      evBlock[e] = s;
}

// Event rows in order: row e of block s is blink perm[e] of the block, with its dye's position and id.
__global__ __launch_bounds__(kThreads) void gather_events_kernel(const int32_t* __restrict__ evBlock, int64_t nEv,
                                                                 const int64_t* __restrict__ dyeStart,
                                                                 const int64_t* __restrict__ winOff,
                                                                 const int64_t* __restrict__ evOff, int32_t nWin,
                                                                 const int32_t* __restrict__ perm,
                                                                 const iscg_blink* __restrict__ blinks,
                                                                 const int32_t* __restrict__ winIdx,
                                                                 const DyeRec* __restrict__ dyes,
                                                                 const BlockMeta* __restrict__ meta, int64_t capacity,
                                                                 iscg_world_event* __restrict__ out)
{
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= nEv || e >= capacity)
      return;
   const int s = evBlock[e];
   int64_t e0, e1;
   seg_of(s, dyeStart, winOff, evOff, e0, e1);
   const int64_t src = e0 + perm[e];
   // the blink's in-window dye: the last w with evOff[w] <= src
   int32_t lo = 0, hi = nWin;
   while (hi - lo > 1)
   {
      const int32_t mid = (lo + hi) >> 1;
      if (evOff[mid] <= src)
         lo = mid;
      else
         hi = mid;
   }
   const iscg_blink b = blinks[src];
   const DyeRec d = dyes[winIdx[lo]];
   iscg_world_event r{};
   r.x = d.x;
   r.y = d.y;
   r.z = d.z;
   r.tOn = b.tOn;
   r.tOff = b.tOff;
   r.brightness = b.brightness;
   r.aux = 0;
   r.id = d.id;
   r.structure = (uint8_t)meta[s].structure;
   r.state = 0;
   out[e] = r;
}

// LabelSchedule's continuous window of each in-window dye (core/src/dyes.cpp:440-478, under the structure's history)
// ending after tMin: 1 if it has one (the window in win[i]).
__global__ __launch_bounds__(kThreads) void cont_kernel(const iscg_dye_address* __restrict__ addr,
                                                        const uint8_t* __restrict__ structure, int32_t n, QueryArgs q,
                                                        double tMin, iscg_cont_window* __restrict__ win,
                                                        int32_t* __restrict__ flag)
{
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n)
      return;
   const int st = structure[i];
   bool one = false;
   iscg_cont_window w{};
   if (st < 3)
   {
      const LabelDev& L = q.lab[st];
      const iscg_dye_address d = addr[i];
      one = cont_window(L.mode, L.preState, L.kin, q.hist[st],
                        Address::of(d.h1, (uint32_t)d.k, (uint32_t)d.n), w);
   }
   one = one && w.tOff > tMin;
   win[i] = w;
   flag[i] = one;
}

__global__ __launch_bounds__(kThreads) void cont_rows_kernel(const iscg_cont_window* __restrict__ win,
                                                             const int32_t* __restrict__ flag,
                                                             const int64_t* __restrict__ off, int32_t n,
                                                             const int32_t* __restrict__ winIdx,
                                                             const uint8_t* __restrict__ structure,
                                                             const DyeRec* __restrict__ dyes, int64_t capacity,
                                                             iscg_world_event* __restrict__ out)
{
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n || !flag[i] || off[i] >= capacity)
      return;
   const DyeRec d = dyes[winIdx[i]];
   const iscg_cont_window w = win[i];
   iscg_world_event r{};
   r.x = d.x;
   r.y = d.y;
   r.z = d.z;
   r.tOn = w.tOn;
   r.tOff = w.tOff;
   r.brightness = 1.0;
   r.aux = w.aux;
   r.id = d.id;
   r.structure = structure[i];
   r.state = w.state;
   out[off[i]] = r;
}

// The most blinks of one selected block (block s's blinks are evOff[winOff[dyeStart[s]]] .. of s + 1): the longest
// block decides the merge passes. One workgroup; the maximum is the same in any order.
__global__ __launch_bounds__(1024) void longest_block_kernel(const int64_t* __restrict__ dyeStart,
                                                             const int64_t* __restrict__ winOff,
                                                             const int64_t* __restrict__ evOff, int64_t nSel,
                                                             int64_t* __restrict__ out)
{
   __shared__ long long part[16];
   long long m = 0;
   for (int64_t s = threadIdx.x; s < nSel; s += 1024)
   {
      const long long n = evOff[winOff[dyeStart[s + 1]]] - evOff[winOff[dyeStart[s]]];
      m = n > m ? n : m;
   }
   for (int d = 32; d > 0; d >>= 1)
   {
      const long long o = __shfl_xor(m, d, 64);
      m = o > m ? o : m;
   }
   if ((threadIdx.x & 63) == 0)
      part[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0)
   {
      long long r = 0;
      for (int i = 0; i < 16; ++i)
         r = part[i] > r ? part[i] : r;
      *out = r;
   }
}

using PersistWin = PersistGen<true, iscg_blink>;
using PersistHistWin = PersistHistGen<true, iscg_blink>;

unsigned blocks_for(int64_t n, int per = kThreads) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---- the world object ----------------------------------------------------------------------------------------------

struct iscg_world
{
   iscg_context* h = nullptr;
   uint32_t seed = 0;
   int32_t nCells = 0, nMts = 0, nCands = 0;
   int64_t nPts = 0;
   bool hasNpcs = false, hasLamin = false;
   double npcSiteReach = 0;
   iscg_label labels[kStructures] = {};
   double protoOff[kProto] = {};
   // World::SetKineticsHistory: the segment starts and each structure's kinetics per segment (host copies), the
   // segments of two or more on the device
   std::vector<double> histT;
   std::vector<iscg_kinetics> histKin[kStructures];
   DevBuf histSeg[kStructures];
   // geometry
   DevBuf cells, mts, pts, cum, T, U, V, cands, npcB, lamB, npcIds, lamIds, npcs, lamin, npcSites, npcStart,
      laminStart, npcReach, laminReach;
   // query scratch
   DevBuf flag, off, part, sel, counts, meta, dyeStart, dyes, dyeBlock, winFlag, winOff, addr, structure, winIdx,
      evCounts, evOff, blinks, perm, perm2, evBlock, contWin, total;
};

namespace {

int32_t wfail(iscg_world* w, const std::string& m) { return fail(w->h, m); }

bool wok(iscg_world* w, hipError_t e, const char* what) { return ok(w->h, e, what); }

template <typename T>
bool upload(iscg_world* w, DevBuf& b, const T* src, size_t n, const char* what)
{
   if (!wok(w, b.reserve((n ? n : 1) * sizeof(T)), what))
      return false;
   return n == 0 || wok(w, hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, w->h->stream), what);
}

template <typename T>
bool reserve(iscg_world* w, DevBuf& b, int64_t n, const char* what)
{
   return wok(w, b.reserve((size_t)(n > 0 ? n : 1) * sizeof(T)), what);
}

// Exclusive scan of n counts into off[0 .. n] (off[n] the total); *total (host) receives it.
bool scan(iscg_world* w, const int32_t* v, int64_t n, int64_t* off, int64_t* total)
{
   const int64_t nPart = n ? (n + kScanTile - 1) / kScanTile : 0;
   if (!reserve<long long>(w, w->part, nPart, "world scan"))
      return false;
   hipStream_t st = w->h->stream;
   if (nPart)
      hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nPart), dim3(kThreads), 0, st, v, n,
                         w->part.as<long long>());
   hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, st, w->part.as<long long>(), (int32_t)nPart, off, n);
   if (nPart)
      hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nPart), dim3(kThreads), 0, st, v, n,
                         w->part.as<long long>(), off);
   if (!wok(w, hipGetLastError(), "world scan launch"))
      return false;
   if (!total)
      return true;
   return wok(w, hipMemcpyAsync(total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st), "world scan total") &&
          wok(w, hipStreamSynchronize(st), "world scan total");
}

// NormLinkage(hasLinkage ? linkage : LegacyLinkage(structure)) (core/src/dyes.cpp:66-83)
void norm_linkage(int structure, const iscg_label& l, double out[4])
{
   double v[4] = {l.binderNm, l.binderFlexNm, l.linkerMinNm, l.linkerMaxNm};
   if (!l.hasLinkage)
   {
      const double mt[4] = {12, 0, 2, 5}, npc[4] = {0, 0, 2, 5}, none[4] = {0, 0, 0, 0};
      std::memcpy(v, structure == 0 ? mt : structure == 1 ? npc : none, sizeof v);
   }
   auto nz = [](double x) { return x > 0 ? x : 0.0; };
   out[0] = nz(v[0]);
   out[1] = nz(v[1]);
   out[2] = nz(v[2]);
   out[3] = std::max(out[2], nz(v[3]));
}

LabelDev label_dev(int structure, const iscg_label& l)
{
   double lk[4];
   norm_linkage(structure, l, lk);
   LabelDev d{};
   d.below = std::min(1.0, std::max(0.0, l.density));
   d.ff = std::min(1.0, std::max(0.0, l.fluorescentFraction));
   d.mode = l.mode;
   d.preState = l.preState;
   d.kin = l.kin;
   d.binderNm = lk[0];
   d.flexNm = lk[1];
   // LinkageSetup (core/src/dyes.cpp:85-91)
   const double lMin = lk[2] * kNm, lMax = lk[3] * kNm;
   d.minL3 = lMin * lMin * lMin;
   d.maxL3 = lMax * lMax * lMax;
   d.flexUm = lk[1] * kNm;
   return d;
}

QueryArgs query_args(const iscg_world* w, const double* win)
{
   QueryArgs q{};
   q.win = Window{win[0], win[1], win[2], win[3], win[4], win[5]};
   q.seed = w->seed;
   for (int s = 0; s < 3; ++s)
   {
      q.lab[s] = label_dev(s, w->labels[s]);
      // HistoryOf(s): one segment is its kinetics from t = 0, two or more the walk over them
      if (w->histT.size() == 1)
         q.lab[s].kin = w->histKin[s][0];
      q.hist[s] = w->histT.size() >= 2 ? HistDev{(int32_t)w->histT.size(), w->histSeg[s].as<HistSeg>()}
                                       : HistDev{0, nullptr};
   }
   double lk[3][4];
   for (int s = 0; s < 3; ++s)
      norm_linkage(s, w->labels[s], lk[s]);
   // MtDyeReach, NpcDyeReach, LaminDyeReach (core/src/world.cpp:225-246), the walk's block reaches (:552-558)
   q.mtReach = (kMtRadiusNm + lk[0][0] + lk[0][1] + lk[0][3]) * 1e-3;
   q.blockReach = kDyeBlockUm / 2 + q.mtReach;
   const double npcDyeReach = w->npcSiteReach + (lk[1][0] + lk[1][1] + lk[1][3]) * 1e-3;
   const double laminDyeReach = (lk[2][0] + lk[2][1] + lk[2][3]) * 1e-3;
   q.npcRb = std::sqrt(3.0) / 2 * kNpcBlockUm + npcDyeReach;
   q.laminRb = std::sqrt(3.0) / 2 * kNpcBlockUm + laminDyeReach;
   const iscg_label& ln = w->labels[1];
   const iscg_label& ll = w->labels[2];
   q.npcOn = w->hasNpcs && ln.density > 0 && ln.fluorescentFraction > 0;
   q.laminOn = w->hasLamin && ll.density > 0 && ll.fluorescentFraction > 0;
   q.npcSiteReach = w->npcSiteReach;
   return q;
}

Geo geo_of(iscg_world* w)
{
   Geo g{};
   g.cells = w->cells.as<iscg_world_cell>();
   g.cands = w->cands.as<Cand>();
   g.mts = w->mts.as<MtDev>();
   g.pts = w->pts.as<Pt>();
   g.cum = w->cum.as<double>();
   g.T = w->T.as<Pt>();
   g.U = w->U.as<Pt>();
   g.V = w->V.as<Pt>();
   g.npcB = w->npcB.as<EnvBlock>();
   g.lamB = w->lamB.as<EnvBlock>();
   g.npcIds = w->npcIds.as<int32_t>();
   g.lamIds = w->lamIds.as<int32_t>();
   g.npcs = w->npcs.as<iscg_npc>();
   g.lamin = w->lamin.as<iscg_lamin_point>();
   g.npcSites = w->npcSites.as<double>();
   g.npcStart = w->npcStart.as<int32_t>();
   g.laminStart = w->laminStart.as<int32_t>();
   for (int k = 0; k < kProto; ++k)
      g.off[k] = w->protoOff[k];
   return g;
}

bool valid_window(const double* win)
{
   if (!win)
      return false;
   for (int i = 0; i < 6; ++i)
      if (std::isnan(win[i]))
         return false;
   return true;
}

// The candidate blocks the window's walk takes, in walk order (w->sel[0 .. nSel)).
bool select_blocks(iscg_world* w, const QueryArgs& q, int64_t& nSel)
{
   hipStream_t st = w->h->stream;
   const int64_t n = w->nCands;
   if (!reserve<int32_t>(w, w->flag, n, "world flags") || !reserve<int64_t>(w, w->off, n + 1, "world offsets") ||
       !reserve<int32_t>(w, w->sel, n, "world selection"))
      return false;
   if (n)
   {
      ProfScope ps(w->h, "world_select");
      hipLaunchKernelGGL(select_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, w->cands.as<Cand>(), (int32_t)n,
                         w->cells.as<iscg_world_cell>(), w->mts.as<MtDev>(), w->npcReach.as<double>(),
                         w->laminReach.as<double>(), q, w->flag.as<int32_t>());
   }
   if (!scan(w, w->flag.as<int32_t>(), n, w->off.as<int64_t>(), &nSel))
      return false;
   if (n)
      hipLaunchKernelGGL(compact_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, w->flag.as<int32_t>(),
                         w->off.as<int64_t>(), n, w->sel.as<int32_t>());
   return wok(w, hipGetLastError(), "world selection");
}

// The dye blocks of a query: select, count, scan, fill. *nSel, *nDyes (host). raw: only select and count
// (iscg_world_block_dyes writes the generators' rows itself).
bool make_dyes(iscg_world* w, const QueryArgs& q, int64_t& nSel, int64_t& nDyes, bool raw = false)
{
   hipStream_t st = w->h->stream;
   if (!select_blocks(w, q, nSel))
      return false;
   if (!reserve<int32_t>(w, w->counts, nSel, "world block counts") ||
       !reserve<BlockMeta>(w, w->meta, nSel, "world block meta") ||
       !reserve<int64_t>(w, w->dyeStart, nSel + 1, "world dye starts"))
      return false;
   const Geo g = geo_of(w);
   if (nSel)
   {
      ProfScope ps(w->h, "world_block_count");
      hipLaunchKernelGGL(block_count_kernel, dim3((unsigned)nSel), dim3(kThreads), 0, st, g, q, w->sel.as<int32_t>(),
                         (int32_t)nSel, w->counts.as<int32_t>(), w->meta.as<BlockMeta>());
   }
   if (!scan(w, w->counts.as<int32_t>(), nSel, w->dyeStart.as<int64_t>(), &nDyes))
      return false;
   if (nDyes >= (1ll << 31))
   {
      wfail(w, "iscg_world: more than 2^31 dyes in the blocks of one query");
      return false;
   }
   if (!raw && (!reserve<DyeRec>(w, w->dyes, nDyes, "world dyes") ||
                !reserve<int32_t>(w, w->dyeBlock, nDyes, "world dyes")))
      return false;
   if (nSel && !raw)
   {
      ProfScope ps(w->h, "world_block_fill");
      hipLaunchKernelGGL(block_fill_kernel<kFillRecords>, dim3((unsigned)nSel), dim3(kThreads), 0, st, g, q,
                         w->sel.as<int32_t>(), w->dyeStart.as<int64_t>(), w->dyes.as<DyeRec>(),
                         w->dyeBlock.as<int32_t>(), nullptr, 0, DensityArgs{});
   }
   return wok(w, hipGetLastError(), "world dye blocks");
}

// The in-window dyes (mask: their structures): flags and offsets over the dyes, the compacted addresses. *nWin.
bool window_dyes(iscg_world* w, const QueryArgs& q, int64_t nDyes, uint32_t mask, int64_t& nWin)
{
   hipStream_t st = w->h->stream;
   if (!reserve<int32_t>(w, w->winFlag, nDyes, "world window flags") ||
       !reserve<int64_t>(w, w->winOff, nDyes + 1, "world window offsets"))
      return false;
   if (nDyes)
      hipLaunchKernelGGL(win_flag_kernel, dim3(blocks_for(nDyes)), dim3(kThreads), 0, st, w->dyes.as<DyeRec>(),
                         w->dyeBlock.as<int32_t>(), w->meta.as<BlockMeta>(), nDyes, q.win, mask,
                         w->winFlag.as<int32_t>());
   if (!scan(w, w->winFlag.as<int32_t>(), nDyes, w->winOff.as<int64_t>(), &nWin))
      return false;
   if (!reserve<iscg_dye_address>(w, w->addr, nWin, "world addresses") ||
       !reserve<uint8_t>(w, w->structure, nWin, "world addresses") ||
       !reserve<int32_t>(w, w->winIdx, nWin, "world addresses"))
      return false;
   if (nDyes)
      hipLaunchKernelGGL(win_list_kernel, dim3(blocks_for(nDyes)), dim3(kThreads), 0, st, w->dyes.as<DyeRec>(),
                         w->dyeBlock.as<int32_t>(), w->meta.as<BlockMeta>(), w->winFlag.as<int32_t>(),
                         w->winOff.as<int64_t>(), nDyes, w->addr.as<iscg_dye_address>(), w->structure.as<uint8_t>(),
                         w->winIdx.as<int32_t>());
   return wok(w, hipGetLastError(), "world window dyes");
}

int32_t query_prologue(iscg_world* w, const double* win, int64_t* total, const char* what)
{
   if (!w)
      return -1;
   if (begin(w->h))
      return -1;
   if (!valid_window(win) || !total)
      return wfail(w, std::string(what) + ": bad argument");
   return 0;
}

bool valid_label(const iscg_label& l)
{
   const iscg_kinetics& k = l.kin;
   auto fin = [](double v) { return std::isfinite(v); };
   if (l.mode < 0 || l.mode > 3 || (l.preState && l.mode != 1))
      return false;
   if (!fin(l.density) || !(l.density >= 0) || !fin(l.fluorescentFraction) || !(l.fluorescentFraction >= 0))
      return false;
   if (!fin(k.activationRatePerSec) || !(k.activationRatePerSec >= 0) || !fin(k.onSec) || !(k.onSec > 0) ||
       !fin(k.offSec) || !(k.offSec >= 0) || !fin(k.bleachProb) || !(k.bleachProb > 0) || !fin(k.photonCV) ||
       !(k.photonCV >= 0) || !fin(k.initialOnSec) || !(k.initialOnSec >= 0))
      return false;
   if (l.hasLinkage && (!fin(l.binderNm) || !fin(l.binderFlexNm) || !fin(l.linkerMinNm) || !fin(l.linkerMaxNm)))
      return false;
   return true;
}

// The DNA-PAINT bin range of a query and the largest rate it draws at (World::PersistFor, core/src/world.cpp:728-740,
// with PersistentMaxOn, core/src/dyes.cpp:628-634): the label's kinetics, or those of every segment of a history.
bool persist_range(const iscg_kinetics& kin, const std::vector<iscg_kinetics>* hist, double t0, double t1,
                   long long& b0, long long& b1, double& maxRate)
{
   double maxOn = 20.0 * kin.onSec;
   bool any = kin.activationRatePerSec > 0;
   maxRate = kin.activationRatePerSec;
   if (hist && !hist->empty())
   {
      maxOn = 0;
      any = false;
      maxRate = 0;
      for (const iscg_kinetics& q : *hist)
      {
         maxOn = std::max(maxOn, 20.0 * q.onSec);
         any = any || q.activationRatePerSec > 0;
         maxRate = std::max(maxRate, q.activationRatePerSec);
      }
   }
   const long long f = (long long)std::floor((t0 - maxOn) / 1.0);
   b0 = f > 0 ? f : 0;
   b1 = (long long)std::floor(t1 / 1.0);
   return any && t1 > t0;
}

}  // namespace

extern "C" {

int32_t iscg_world_create(iscg_context* h, uint32_t seed, const iscg_world_geometry* g, iscg_world** out)
{
   if (begin(h))
      return -1;
   if (!g || !out || g->nCells < 0 || (g->nCells > 0 && (!g->cells || !g->mtStart)))
      return fail(h, "iscg_world_create: bad argument");
   const int32_t nCells = g->nCells;
   const int32_t nMts = nCells ? g->mtStart[nCells] : 0;
   if (nCells && g->mtStart[0] != 0)
      return fail(h, "iscg_world_create: mtStart[0] must be 0");
   for (int32_t c = 0; c < nCells; ++c)
      if (g->mtStart[c + 1] < g->mtStart[c])
         return fail(h, "iscg_world_create: mtStart must not decrease");
   if (nMts > 0 && (!g->ptStart || !g->pts))
      return fail(h, "iscg_world_create: microtubules without points");
   const int64_t nPts = nMts ? g->ptStart[nMts] : 0;
   if (nMts && g->ptStart[0] != 0)
      return fail(h, "iscg_world_create: ptStart[0] must be 0");
   for (int32_t m = 0; m < nMts; ++m)
      if (g->ptStart[m + 1] < g->ptStart[m] || g->ptStart[m + 1] - g->ptStart[m] > (1 << 30))
         return fail(h, "iscg_world_create: bad ptStart");
   auto env_ok = [&](const int32_t* start, const void* arr) {
      if (!start)
         return true;
      if (start[0] != 0)
         return false;
      for (int32_t c = 0; c < nCells; ++c)
         if (start[c + 1] < start[c])
            return false;
      return start[nCells] == 0 || arr != nullptr;
   };
   if (!env_ok(g->npcStart, g->npcs) || !env_ok(g->laminStart, g->lamin))
      return fail(h, "iscg_world_create: bad pore or lamina ranges");

   iscg_world* w = new iscg_world();
   w->h = h;
   w->seed = seed;
   w->nCells = nCells;
   w->nMts = nMts;
   w->nPts = nPts;
   w->npcSiteReach = g->npcSiteReachUm;
   for (int s = 0; s < kStructures; ++s)
   {
      iscg_label& l = w->labels[s];  // isc::Label's defaults (core/src/dyes.h:129-176)
      l.density = 0.7;
      l.fluorescentFraction = 1.0;
      l.mode = 2;
      l.preState = 0;
      l.kin = iscg_kinetics{0.01, 0.05, 1.0, 1.0, 0.5, 0.0};
      l.hasLinkage = 0;
   }
   // MtProtofilamentOffsetNm (core/src/dyes.cpp:31-34)
   for (int k = 0; k < kProto; ++k)
      w->protoOff[k] = std::fmod(k * kLatticeStart * kDimerNm / kProto, kDimerNm);
   auto bail = [&](iscg_world* ww) {
      delete ww;
      return -1;
   };

   // microtubules
   std::vector<MtDev> mts((size_t)nMts);
   for (int32_t c = 0; c < nCells; ++c)
      for (int32_t m = g->mtStart[c]; m < g->mtStart[c + 1]; ++m)
      {
         MtDev& d = mts[(size_t)m];
         d.p0 = g->ptStart[m];
         d.nPts = (int32_t)(g->ptStart[m + 1] - g->ptStart[m]);
         d.cell = c;
         d.index = g->mtIndex ? g->mtIndex[m] : m - g->mtStart[c];
      }
   if (!upload(w, w->cells, g->cells, (size_t)nCells, "world cells") ||
       !upload(w, w->mts, mts.data(), mts.size(), "world microtubules") ||
       !upload(w, w->pts, reinterpret_cast<const Pt*>(g->pts), (size_t)nPts, "world points") ||
       !reserve<double>(w, w->cum, nPts, "world frames") || !reserve<Pt>(w, w->T, nPts, "world frames") ||
       !reserve<Pt>(w, w->U, nPts, "world frames") || !reserve<Pt>(w, w->V, nPts, "world frames"))
      return bail(w);
   hipStream_t st = h->stream;
   if (nMts)
      hipLaunchKernelGGL(frames_kernel, dim3(blocks_for(nMts)), dim3(kThreads), 0, st, w->pts.as<Pt>(),
                         w->mts.as<MtDev>(), nMts, w->cum.as<double>(), w->T.as<Pt>(), w->U.as<Pt>(), w->V.as<Pt>());
   if (!wok(w, hipGetLastError(), "world frames") ||
       (nMts && !wok(w, hipMemcpyAsync(mts.data(), w->mts.p, mts.size() * sizeof(MtDev), hipMemcpyDeviceToHost, st),
                     "world frames")) ||
       !wok(w, hipStreamSynchronize(st), "world frames"))
      return bail(w);

   // envelope blocks (EnvelopeBlocks, core/src/world.cpp:99-130): per cell, the points per cube, cubes in (iz, iy, ix)
   // order, points in candidate order
   std::vector<EnvBlock> envB[2];
   std::vector<int32_t> envIds[2];
   std::vector<std::vector<int32_t>> envOfCell[2];
   const int32_t* starts[2] = {g->npcStart, g->laminStart};
   for (int e = 0; e < 2; ++e)
   {
      envOfCell[e].resize((size_t)nCells);
      if (!starts[e])
         continue;
      for (int32_t c = 0; c < nCells; ++c)
      {
         std::map<std::array<int32_t, 3>, std::vector<int32_t>> cubes;
         for (int32_t i = starts[e][c]; i < starts[e][c + 1]; ++i)
         {
            const double* q = e == 0 ? g->npcs[i].c : g->lamin[i].c;
            const int32_t ix = (int32_t)std::floor(q[0] / kNpcBlockUm), iy = (int32_t)std::floor(q[1] / kNpcBlockUm);
            const int32_t iz = (int32_t)std::floor(q[2] / kNpcBlockUm);
            cubes[{iz, iy, ix}].push_back(i);
         }
         for (const auto& kv : cubes)
         {
            EnvBlock b{};
            b.cell = c;
            b.first = (int32_t)envIds[e].size();
            b.count = (int32_t)kv.second.size();
            b.x = (kv.first[2] + 0.5) * kNpcBlockUm;
            b.y = (kv.first[1] + 0.5) * kNpcBlockUm;
            b.z = (kv.first[0] + 0.5) * kNpcBlockUm;
            envOfCell[e][(size_t)c].push_back((int32_t)envB[e].size());
            envB[e].push_back(b);
            envIds[e].insert(envIds[e].end(), kv.second.begin(), kv.second.end());
         }
      }
   }
   w->hasNpcs = !envB[0].empty();
   w->hasLamin = !envB[1].empty();

   // the candidate blocks in the walk's order (ForEachDyeBlock, core/src/world.cpp:556-589)
   std::vector<Cand> cands;
   for (int32_t c = 0; c < nCells; ++c)
   {
      for (int32_t m = g->mtStart[c]; m < g->mtStart[c + 1]; ++m)
      {
         const MtDev& d = mts[(size_t)m];
         if (d.nPts < 2)
            continue;
         const int nBlocks = (int)std::ceil(d.len / kDyeBlockUm);
         for (int b = 0; b < nBlocks; ++b)
            cands.push_back(Cand{c, 0, m, b, 0, 0, 0});
      }
      for (int e = 0; e < 2; ++e)
         for (size_t b = 0; b < envOfCell[e][(size_t)c].size(); ++b)
            cands.push_back(Cand{c, 1 + e, envOfCell[e][(size_t)c][b], (int32_t)b, 0, 0, 0});
   }
   if (cands.size() >= (size_t)(1 << 30))
   {
      fail(h, "iscg_world_create: too many dye blocks");
      return bail(w);
   }
   w->nCands = (int32_t)cands.size();
   const int32_t nNpc = g->npcStart && nCells ? g->npcStart[nCells] : 0;
   const int32_t nLam = g->laminStart && nCells ? g->laminStart[nCells] : 0;
   std::vector<int32_t> zeros((size_t)nCells + 1, 0);
   if (!upload(w, w->cands, cands.data(), cands.size(), "world blocks") ||
       !upload(w, w->npcB, envB[0].data(), envB[0].size(), "world pore blocks") ||
       !upload(w, w->lamB, envB[1].data(), envB[1].size(), "world lamina blocks") ||
       !upload(w, w->npcIds, envIds[0].data(), envIds[0].size(), "world pore blocks") ||
       !upload(w, w->lamIds, envIds[1].data(), envIds[1].size(), "world lamina blocks") ||
       !upload(w, w->npcs, g->npcs, (size_t)nNpc, "world pores") ||
       !upload(w, w->lamin, g->lamin, (size_t)nLam, "world lamina") ||
       !upload(w, w->npcSites, g->npcSites, 192, "world pore sites") ||
       !upload(w, w->npcStart, g->npcStart ? g->npcStart : zeros.data(), (size_t)nCells + 1, "world pores") ||
       !upload(w, w->laminStart, g->laminStart ? g->laminStart : zeros.data(), (size_t)nCells + 1, "world lamina") ||
       !reserve<double>(w, w->npcReach, nCells, "world reach") ||
       !reserve<double>(w, w->laminReach, nCells, "world reach"))
      return bail(w);
   if (w->nCands)
      hipLaunchKernelGGL(mids_kernel, dim3(blocks_for(w->nCands)), dim3(kThreads), 0, st, w->cands.as<Cand>(),
                         w->nCands, w->cells.as<iscg_world_cell>(), w->mts.as<MtDev>(), w->pts.as<Pt>(),
                         w->cum.as<double>(), w->T.as<Pt>(), w->npcB.as<EnvBlock>(), w->lamB.as<EnvBlock>());
   if (nCells)
   {
      hipLaunchKernelGGL(env_reach_kernel, dim3(nCells), dim3(kThreads), 0, st, w->npcStart.as<int32_t>(),
                         reinterpret_cast<const Pt*>(w->npcs.as<iscg_npc>()), (int)(sizeof(iscg_npc) / sizeof(double)),
                         nCells, w->npcReach.as<double>());
      hipLaunchKernelGGL(env_reach_kernel, dim3(nCells), dim3(kThreads), 0, st, w->laminStart.as<int32_t>(),
                         reinterpret_cast<const Pt*>(w->lamin.as<iscg_lamin_point>()),
                         (int)(sizeof(iscg_lamin_point) / sizeof(double)), nCells, w->laminReach.as<double>());
   }
   if (!wok(w, hipGetLastError(), "world blocks") || !wok(w, hipStreamSynchronize(st), "world blocks"))
      return bail(w);
   *out = w;
   return 0;
}

int32_t iscg_world_destroy(iscg_world* w)
{
   if (!w)
      return 0;
   (void)hipStreamSynchronize(w->h->stream);
   delete w;
   return 0;
}

int32_t iscg_world_set_label(iscg_world* w, int32_t structure, const iscg_label* label)
{
   if (!w)
      return -1;
   if (begin(w->h))
      return -1;
   if (structure < 0 || structure >= kStructures || !label || !valid_label(*label))
      return wfail(w, "iscg_world_set_label: bad label");
   w->labels[structure] = *label;
   return 0;
}

int32_t iscg_world_set_history(iscg_world* w, const double* tStart, const iscg_kinetics* kin, int32_t n)
{
   if (!w)
      return -1;
   if (begin(w->h))
      return -1;
   if (n < 0 || n > (1 << 16) || (n > 0 && (!tStart || !kin)))
      return wfail(w, "iscg_world_set_history: bad argument");
   // World::SetKineticsHistory's checks (core/src/world.cpp:933-945)
   for (int32_t i = 0; i < n; ++i)
      if (!std::isfinite(tStart[i]) || (i == 0 ? tStart[i] != 0 : !(tStart[i] > tStart[i - 1])))
         return wfail(w, "iscg_world_set_history: segment starts must be 0, then finite and increasing");
   for (int64_t i = 0; i < (int64_t)n * kStructures; ++i)
   {
      iscg_label l = w->labels[0];
      l.kin = kin[i];
      if (!valid_label(l))
         return wfail(w, "iscg_world_set_history: bad kinetics");
   }
   std::vector<HistSeg> segs((size_t)n);
   for (int s = 0; s < kStructures; ++s)
   {
      w->histKin[s].resize((size_t)n);
      for (int32_t i = 0; i < n; ++i)
      {
         w->histKin[s][(size_t)i] = kin[(size_t)i * kStructures + (size_t)s];
         segs[(size_t)i] = HistSeg{tStart[i], w->histKin[s][(size_t)i]};
      }
      if (n >= 2 && (!upload(w, w->histSeg[s], segs.data(), segs.size(), "iscg_world_set_history") ||
                     !wok(w, hipStreamSynchronize(w->h->stream), "iscg_world_set_history")))
      {
         w->histT.clear();
         return -1;
      }
   }
   w->histT.assign(tStart, tStart + n);
   return 0;
}

int32_t iscg_world_block_dyes(iscg_world* w, const double window[6], iscg_block_dye* out, int64_t capacity,
                              int64_t* total)
{
   if (query_prologue(w, window, total, "iscg_world_block_dyes"))
      return -1;
   if (capacity < 0 || (capacity > 0 && !out))
      return wfail(w, "iscg_world_block_dyes: bad argument");
   const QueryArgs q = query_args(w, window);
   int64_t nSel = 0, nDyes = 0;
   if (!make_dyes(w, q, nSel, nDyes, true))
      return -1;
   *total = nDyes;
   const int64_t nOut = std::min(nDyes, capacity);
   if (nOut == 0)
      return 0;
   iscg_block_dye* dout = (iscg_block_dye*)stage_out(w->h, 0, out, (size_t)nOut * sizeof(iscg_block_dye));
   if (!dout)
      return -1;
   const Geo g = geo_of(w);
   hipLaunchKernelGGL(block_fill_kernel<kFillRaw>, dim3((unsigned)nSel), dim3(kThreads), 0, w->h->stream, g, q,
                      w->sel.as<int32_t>(), w->dyeStart.as<int64_t>(), nullptr, nullptr, dout, nOut, DensityArgs{});
   if (!wok(w, hipGetLastError(), "iscg_world_block_dyes"))
      return -1;
   return finish_out(w->h, out, dout, (size_t)nOut * sizeof(iscg_block_dye)) ? 0 : -1;
}

int32_t iscg_world_sites(iscg_world* w, const double window[6], iscg_world_dye* out, int64_t capacity,
                         int64_t* total)
{
   if (query_prologue(w, window, total, "iscg_world_sites"))
      return -1;
   if (capacity < 0 || (capacity > 0 && !out))
      return wfail(w, "iscg_world_sites: bad argument");
   const QueryArgs q = query_args(w, window);
   int64_t nSel = 0, nDyes = 0, nWin = 0;
   if (!make_dyes(w, q, nSel, nDyes) || !window_dyes(w, q, nDyes, 0xFu, nWin))
      return -1;
   *total = nWin;
   const int64_t nOut = std::min(nWin, capacity);
   if (nOut == 0)
      return 0;
   iscg_world_dye* dout = (iscg_world_dye*)stage_out(w->h, 0, out, (size_t)nOut * sizeof(iscg_world_dye));
   if (!dout)
      return -1;
   hipLaunchKernelGGL(sites_kernel, dim3(blocks_for(nDyes)), dim3(kThreads), 0, w->h->stream, w->dyes.as<DyeRec>(),
                      w->dyeBlock.as<int32_t>(), w->meta.as<BlockMeta>(), w->cells.as<iscg_world_cell>(),
                      w->winFlag.as<int32_t>(), w->winOff.as<int64_t>(), nDyes, nOut, dout);
   if (!wok(w, hipGetLastError(), "iscg_world_sites"))
      return -1;
   return finish_out(w->h, out, dout, (size_t)nOut * sizeof(iscg_world_dye)) ? 0 : -1;
}

int32_t iscg_world_density3d(iscg_world* w, const double window[6], int32_t nx, int32_t ny, int32_t nz,
                             uint32_t structureMask, float* out, int64_t* total)
{
   if (query_prologue(w, window, total, "iscg_world_density3d"))
      return -1;
   if (nx <= 0 || ny <= 0 || nz <= 0 || !out || (int64_t)nx * ny * nz > (1ll << 34))
      return wfail(w, "iscg_world_density3d: bad argument");
   if (!(std::isfinite(window[0]) && std::isfinite(window[1]) && std::isfinite(window[2]) &&
         std::isfinite(window[3])))
      return wfail(w, "iscg_world_density3d: x0, y0, x1, y1 must be finite");
   if (nz > 1 && !(std::isfinite(window[4]) && std::isfinite(window[5])))
      return wfail(w, "iscg_world_density3d: zMin and zMax must be finite with nz > 1");
   const int64_t nVox = (int64_t)nx * ny * nz;
   float* dout = (float*)stage_out(w->h, 0, out, (size_t)nVox * sizeof(float));
   if (!dout || !reserve<unsigned long long>(w, w->total, 1, "iscg_world_density3d"))
      return -1;
   hipStream_t st = w->h->stream;
   hipLaunchKernelGGL(zero_kernel, dim3(grid_for((size_t)nVox, kThreads)), dim3(kThreads), 0, st, dout, nVox,
                      w->total.as<unsigned long long>());
   unsigned long long tot = 0;
   if (structureMask)
   {
      const QueryArgs q = query_args(w, window);
      int64_t nSel = 0;
      if (!select_blocks(w, q, nSel))
         return -1;
      // Density3dInWindow's scales (core/src/world.cpp:1527-1529)
      DensityArgs den{};
      den.out = dout;
      den.total = w->total.as<unsigned long long>();
      den.nx = nx;
      den.ny = ny;
      den.nz = nz;
      den.mask = structureMask;
      den.sx = nx / (window[2] - window[0]);
      den.sy = ny / (window[3] - window[1]);
      den.sz = nz > 1 ? nz / (window[5] - window[4]) : 0.0;
      if (nSel)
      {
         ProfScope ps(w->h, "world_density");
         hipLaunchKernelGGL(block_fill_kernel<kFillDensity>, dim3((unsigned)nSel), dim3(kThreads), 0, st, geo_of(w), q,
                            w->sel.as<int32_t>(), nullptr, nullptr, nullptr, nullptr, 0, den);
      }
   }
   if (!wok(w, hipGetLastError(), "iscg_world_density3d") ||
       !wok(w, hipMemcpyAsync(&tot, w->total.p, sizeof tot, hipMemcpyDeviceToHost, st), "iscg_world_density3d") ||
       !wok(w, hipStreamSynchronize(st), "iscg_world_density3d"))
      return -1;
   *total = (int64_t)tot;
   return finish_out(w->h, out, dout, (size_t)nVox * sizeof(float)) ? 0 : -1;
}

int32_t iscg_world_continuous(iscg_world* w, const double window[6], double tMin, iscg_world_event* out,
                              int64_t capacity, int64_t* total)
{
   if (query_prologue(w, window, total, "iscg_world_continuous"))
      return -1;
   if (capacity < 0 || (capacity > 0 && !out) || std::isnan(tMin))
      return wfail(w, "iscg_world_continuous: bad argument");
   const QueryArgs q = query_args(w, window);
   int64_t nSel = 0, nDyes = 0, nWin = 0, nCont = 0;
   if (!make_dyes(w, q, nSel, nDyes) || !window_dyes(w, q, nDyes, 0xFu, nWin))
      return -1;
   hipStream_t st = w->h->stream;
   if (!reserve<iscg_cont_window>(w, w->contWin, nWin, "iscg_world_continuous") ||
       !reserve<int32_t>(w, w->evCounts, nWin, "iscg_world_continuous") ||
       !reserve<int64_t>(w, w->evOff, nWin + 1, "iscg_world_continuous"))
      return -1;
   if (nWin)
      hipLaunchKernelGGL(cont_kernel, dim3(blocks_for(nWin)), dim3(kThreads), 0, st,
                         w->addr.as<iscg_dye_address>(), w->structure.as<uint8_t>(), (int32_t)nWin, q, tMin,
                         w->contWin.as<iscg_cont_window>(), w->evCounts.as<int32_t>());
   if (!scan(w, w->evCounts.as<int32_t>(), nWin, w->evOff.as<int64_t>(), &nCont))
      return -1;
   *total = nCont;
   const int64_t nOut = std::min(nCont, capacity);
   if (nOut == 0)
      return 0;
   iscg_world_event* dout = (iscg_world_event*)stage_out(w->h, 0, out, (size_t)nOut * sizeof(iscg_world_event));
   if (!dout)
      return -1;
   hipLaunchKernelGGL(cont_rows_kernel, dim3(blocks_for(nWin)), dim3(kThreads), 0, st,
                      w->contWin.as<iscg_cont_window>(), w->evCounts.as<int32_t>(), w->evOff.as<int64_t>(),
                      (int32_t)nWin, w->winIdx.as<int32_t>(), w->structure.as<uint8_t>(), w->dyes.as<DyeRec>(), nOut,
                      dout);
   if (!wok(w, hipGetLastError(), "iscg_world_continuous"))
      return -1;
   return finish_out(w->h, out, dout, (size_t)nOut * sizeof(iscg_world_event)) ? 0 : -1;
}

int32_t iscg_world_events(iscg_world* w, const double window[6], double t0, double t1, iscg_world_event* out,
                          int64_t capacity, int64_t* total)
{
   if (query_prologue(w, window, total, "iscg_world_events"))
      return -1;
   if (capacity < 0 || (capacity > 0 && !out) || std::isnan(t0) || std::isnan(t1))
      return wfail(w, "iscg_world_events: bad argument");
   const QueryArgs q = query_args(w, window);
   // the schedules of each structure (World::Schedule / PersistentCover, core/src/world.cpp:979-1078)
   long long pb0[3] = {0, 0, 0}, pb1[3] = {-1, -1, -1};
   bool persist[3] = {false, false, false};
   for (int s = 0; s < 3; ++s)
      if (q.lab[s].mode == 2)
      {
         double maxRate = 0;
         persist[s] = persist_range(w->labels[s].kin, w->histT.empty() ? nullptr : &w->histKin[s], t0, t1, pb0[s],
                                    pb1[s], maxRate);
         const long long lo = pb0[s] > 0 ? pb0[s] : 0;
         if (persist[s] && (pb1[s] - lo >= (1ll << 20) || maxRate > 1e5))
            return wfail(w, "iscg_world_events: more than 2^20 time bins or 1e5 blinks per bin for DNA-PAINT");
      }
   // World::Schedule's horizon (twice the query's end); the blinks starting at or after t1 are not wanted
   const double horizon = 2 * t1;
   int64_t nSel = 0, nDyes = 0, nWin = 0, nEv = 0;
   if (!make_dyes(w, q, nSel, nDyes) || !window_dyes(w, q, nDyes, 0xFu, nWin))
      return -1;
   if (nWin >= (1ll << 31))
      return wfail(w, "iscg_world_events: more than 2^31 dyes in the window");
   hipStream_t st = w->h->stream;
   if (!reserve<int32_t>(w, w->evCounts, nWin, "iscg_world_events") ||
       !reserve<int64_t>(w, w->evOff, nWin + 1, "iscg_world_events"))
      return -1;
   const int32_t nw = (int32_t)nWin;
   const unsigned grid = nw ? (unsigned)((nw + kWgDyes - 1) / kWgDyes) : 0;
   if (nw)
      hipLaunchKernelGGL(zero_counts_kernel, dim3(grid_for((size_t)nw, kThreads)), dim3(kThreads), 0, st,
                         w->evCounts.as<int32_t>(), (int64_t)nw);
   // count per structure
   for (int s = 0; s < 3 && nw; ++s)
   {
      const OfStructure sel{w->structure.as<uint8_t>(), (uint8_t)s};
      ProfScope ps(w->h, "world_schedule_count");
      const bool multi = q.hist[s].n >= 2;
      if ((q.lab[s].mode == 0 || q.lab[s].mode == 1) && multi)
      {
         const LabelHistGen g{q.lab[s].mode, q.hist[s], horizon, t0, t1};
         hipLaunchKernelGGL(world_count_kernel<LabelHistGen>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), sel);
      }
      else if (q.lab[s].mode == 0 || q.lab[s].mode == 1)
      {
         const LabelGen g{q.lab[s].mode, q.lab[s].kin, horizon, t0, t1};
         hipLaunchKernelGGL(world_count_kernel<LabelGen>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), sel);
      }
      else if (q.lab[s].mode == 2 && persist[s] && multi)
      {
         const PersistHistWin g{q.hist[s], t0, t1, pb0[s], pb1[s]};
         hipLaunchKernelGGL(world_count_kernel<PersistHistWin>, dim3(grid), dim3(kDyeThreads), 0, st,
                            g, w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), sel);
      }
      else if (q.lab[s].mode == 2 && persist[s])
      {
         const PersistWin g{q.lab[s].kin, t0, t1, pb0[s], pb1[s]};
         hipLaunchKernelGGL(world_count_kernel<PersistWin>, dim3(grid), dim3(kDyeThreads), 0, st,
                            g, w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), sel);
      }
   }
   if (!scan(w, w->evCounts.as<int32_t>(), nWin, w->evOff.as<int64_t>(), &nEv))
      return -1;
   if (nEv >= (1ll << 31))
      return wfail(w, "iscg_world_events: more than 2^31 blinks in one query");
   *total = nEv;
   if (!reserve<iscg_blink>(w, w->blinks, nEv, "iscg_world_events") ||
       !reserve<int32_t>(w, w->perm, nEv, "iscg_world_events") ||
       !reserve<int32_t>(w, w->evBlock, nEv, "iscg_world_events"))
      return -1;
   for (int s = 0; s < 3 && nw && nEv; ++s)
   {
      const OfStructure sel{w->structure.as<uint8_t>(), (uint8_t)s};
      ProfScope ps(w->h, "world_schedule_fill");
      const bool multi = q.hist[s].n >= 2;
      if ((q.lab[s].mode == 0 || q.lab[s].mode == 1) && multi)
      {
         const LabelHistGen g{q.lab[s].mode, q.hist[s], horizon, t0, t1};
         hipLaunchKernelGGL(world_fill_kernel<LabelHistGen>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), w->evOff.as<int64_t>(),
                            w->blinks.as<iscg_blink>(), nEv, sel);
      }
      else if (q.lab[s].mode == 0 || q.lab[s].mode == 1)
      {
         const LabelGen g{q.lab[s].mode, q.lab[s].kin, horizon, t0, t1};
         hipLaunchKernelGGL(world_fill_kernel<LabelGen>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), w->evOff.as<int64_t>(),
                            w->blinks.as<iscg_blink>(), nEv, sel);
      }
      else if (q.lab[s].mode == 2 && persist[s] && multi)
      {
         const PersistHistWin g{q.hist[s], t0, t1, pb0[s], pb1[s]};
         hipLaunchKernelGGL(world_fill_kernel<PersistHistWin>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), w->evOff.as<int64_t>(),
                            w->blinks.as<iscg_blink>(), nEv, sel);
      }
      else if (q.lab[s].mode == 2 && persist[s])
      {
         const PersistWin g{q.lab[s].kin, t0, t1, pb0[s], pb1[s]};
         hipLaunchKernelGGL(world_fill_kernel<PersistWin>, dim3(grid), dim3(kDyeThreads), 0, st, g,
                            w->addr.as<iscg_dye_address>(), nw, w->evCounts.as<int32_t>(), w->evOff.as<int64_t>(),
                            w->blinks.as<iscg_blink>(), nEv, sel);
      }
   }
   const int64_t nOut = std::min(nEv, capacity);
   if (nEv == 0 || nOut == 0)
      return wok(w, hipGetLastError(), "iscg_world_events") ? 0 : -1;
   // per block: the order of its blinks; the longest block decides the merge passes
   hipLaunchKernelGGL(ev_block_kernel, dim3(blocks_for(nw)), dim3(kThreads), 0, st, w->evOff.as<int64_t>(), nw,
                      w->winIdx.as<int32_t>(), w->dyeBlock.as<int32_t>(), w->evBlock.as<int32_t>());
   {
      ProfScope ps(w->h, "world_sort");
      hipLaunchKernelGGL(seg_sort_kernel, dim3((unsigned)nSel), dim3(kSortThreads), 0, st, w->meta.as<BlockMeta>(), q,
                         w->dyeStart.as<int64_t>(), w->winOff.as<int64_t>(), w->evOff.as<int64_t>(),
                         w->blinks.as<iscg_blink>(), w->perm.as<int32_t>());
   }
   // the longest block, reduced on the device (one value back, not the offsets)
   if (!reserve<int64_t>(w, w->total, 1, "iscg_world_events"))
      return -1;
   hipLaunchKernelGGL(longest_block_kernel, dim3(1), dim3(1024), 0, st, w->dyeStart.as<int64_t>(),
                      w->winOff.as<int64_t>(), w->evOff.as<int64_t>(), nSel, w->total.as<int64_t>());
   int64_t longest = 0;
   if (!wok(w, hipGetLastError(), "iscg_world_events") ||
       !wok(w, hipMemcpyAsync(&longest, w->total.p, sizeof longest, hipMemcpyDeviceToHost, st), "iscg_world_events") ||
       !wok(w, hipStreamSynchronize(st), "iscg_world_events"))
      return -1;
   int32_t* cur = w->perm.as<int32_t>();
   if (longest > kSortMax)
   {
      if (!reserve<int32_t>(w, w->perm2, nEv, "iscg_world_events"))
         return -1;
      int32_t* nxt = w->perm2.as<int32_t>();
      ProfScope ps(w->h, "world_merge");
      for (int64_t R = kSortMax; R < longest; R *= 2)
      {
         hipLaunchKernelGGL(merge_pass_kernel, dim3(blocks_for(nEv)), dim3(kThreads), 0, st,
                            w->evBlock.as<int32_t>(), nEv, w->meta.as<BlockMeta>(), q, w->dyeStart.as<int64_t>(),
                            w->winOff.as<int64_t>(), w->evOff.as<int64_t>(), w->blinks.as<iscg_blink>(), (int32_t)R,
                            cur, nxt);
         std::swap(cur, nxt);
      }
   }
   iscg_world_event* dout = (iscg_world_event*)stage_out(w->h, 0, out, (size_t)nOut * sizeof(iscg_world_event));
   if (!dout)
      return -1;
   {
      ProfScope ps(w->h, "world_gather");
      hipLaunchKernelGGL(gather_events_kernel, dim3(blocks_for(nOut)), dim3(kThreads), 0, st,
                         w->evBlock.as<int32_t>(), nEv, w->dyeStart.as<int64_t>(), w->winOff.as<int64_t>(),
                         w->evOff.as<int64_t>(), nw, cur, w->blinks.as<iscg_blink>(), w->winIdx.as<int32_t>(),
                         w->dyes.as<DyeRec>(), w->meta.as<BlockMeta>(), nOut, dout);
   }
   if (!wok(w, hipGetLastError(), "iscg_world_events"))
      return -1;
   return finish_out(w->h, out, dout, (size_t)nOut * sizeof(iscg_world_event)) ? 0 : -1;
}

}  // extern "C"
