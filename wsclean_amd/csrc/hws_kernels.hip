// Device kernels of the MI355X w-stacking core and their launchers.
//
// Invert (gridding/wstackinggridder.cpp:127-137, 323-336, 517-597, 691-715,
// 773-827, 851-907) and predict (cpp:139-157, 599-689, 829-849, 909-969) in
// two forms:
//
// Batched form (HWS_PRECISION_REFERENCE / _MIXED):
//   1. binning: every (row, channel) sample is placed, in FP64 exactly as
//      AddDataSample does, onto its w-layer and uv position. Its K x K
//      footprint touches one to four T x T uv tiles; a record (8 bytes) is
//      emitted per touched (layer, tile) bin, the one holding the footprint
//      start marked primary. A counting sort (histogram, exclusive scan,
//      scatter, with one atomic per run of same-bin lanes) orders records by
//      bin.
//   2. grid (owner computes): a workgroup owns a (layer, tile), deposits all
//      its records into an LDS tile (LDS float atomics, one lane per tap,
//      register pre-accumulation while the footprint stays put) and writes the
//      tile to the layer once with plain coalesced stores; empty tiles are
//      written as zeros. No global atomics and no separate layer clearing.
//   3. degrid mirrors it: the (T + K - 1)^2 subgrid of a (layer, tile) is
//      gathered into LDS once, then one lane per primary record sums its K*K
//      taps.
//   Image-domain passes loop over all layers of a pass per pixel: the
//   w-correction phase is computed once per pixel and layer, and the layer
//   sum is Kahan-compensated in FP32.
//
// Plain form (HWS_PRECISION_PLAIN_FP32): one launch per channel and image,
// FP32 index and phase arithmetic, a global atomic per tap, one projection
// launch per layer.
#include "hws_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hws {

hws_status setLastError(hws_gridder* g, hws_status s, const std::string& msg) {
  if (g)
    g->err = msg;
  else
    threadError() = msg;
  return s;
}

hws_status checkHip(hws_gridder* g, hipError_t e, const char* what) {
  if (e == hipSuccess) return HWS_OK;
  const hws_status s =
      (e == hipErrorOutOfMemory) ? HWS_ERR_OUT_OF_MEMORY : HWS_ERR_HIP;
  return setLastError(g, s, std::string(what) + ": " + hipGetErrorString(e));
}

LayerMap layerMap(const hws_gridder* g, int kind) {
  LayerMap m;
  m.nW = g->nW;
  m.minW = g->minW;
  m.maxW = g->maxW;
  m.isComplex = kind;
  return m;
}

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kThreads = 256;

#define HWS_TRY(expr)                  \
  do {                                 \
    hws_status _s = (expr);            \
    if (_s != HWS_OK) return _s;       \
  } while (0)

struct ImgList {
  int n;
  int idx[kMaxImages];
};

struct GeomArgs {
  int W, H, K, OS, mid;
  double dl, dm;
  int tileSize, tilesX, tilesY;
  int tileShift;     // log2(tileSize) when it is a power of two, else -1
  int rotate;        // apply the l/m shift phase rotation to samples
  int fastPhase;     // rotation phasor from an FP64-reduced FP32 turnSinCos
  double lShift, mShift;
  int blocked;       // layer storage layout (cellIndex)
};

// ----------------------------------------------------------------- fill
__global__ void fillKernel(uint32_t* p, uint32_t value, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = value;
}

// Zeroes n 16-byte words (the layers of a pass before gridding).
__global__ void zero16Kernel(uint4* p, size_t n) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = z;
}

// ------------------------------------------------------------- binning
// Layer index as WToLayer computes it (gridding/wstackinggridder.h:182-193);
// returns -1 when the rounded value is negative or >= nW.
__device__ __forceinline__ int layerOf(const LayerMap& m, double w) {
  if (m.nW == 1) return 0;
  double r;
  if (m.isComplex)
    r = round((w + m.maxW) * (m.nW - 1) / (m.maxW + m.maxW));
  else
    r = round((fabs(w) - m.minW) * (m.nW - 1) / (m.maxW - m.minW));
  if (!(r >= 0.0) || r >= double(m.nW)) return -1;
  return int(r);
}

// Placement of one sample: the arithmetic of AddDataSample
// (gridding/wstackinggridder.cpp:528-564), FP64 throughout.
struct Placement {
  int layer;   // -1 off layer
  int onGrid;  // 0 rejected by the acceptance window
  int x, y;    // wrapped grid position of the footprint centre
  int kx, ky;  // kernel table indices
  int conj;
};

__device__ __forceinline__ Placement place(const GeomArgs& ga,
                                           const LayerMap& lm, double u,
                                           double v, double w) {
  Placement p;
  p.conj = 0;
  if (w < 0.0 && !lm.isComplex) {
    u = -u;
    v = -v;
    w = -w;
    p.conj = 1;
  }
  p.layer = layerOf(lm, w);
  const double xExact = u * ga.dl * double(ga.W);
  const double yExact = v * ga.dm * double(ga.H);
  p.onGrid = 0;
  p.x = p.y = p.kx = p.ky = 0;
  // bound before the int conversion (the reference converts unchecked)
  if (!(fabs(xExact) < 1.0e9) || !(fabs(yExact) < 1.0e9)) return p;
  int x = int(round(xExact));
  int y = int(round(yExact));
  int kx = int(round((xExact - double(x)) * ga.OS));
  int ky = int(round((yExact - double(y)) * ga.OS));
  kx = (kx + (ga.OS * 3) / 2) % ga.OS;
  ky = (ky + (ga.OS * 3) / 2) % ga.OS;
  if (x > -ga.W / 2 && y > -ga.H / 2 && x <= ga.W / 2 && y <= ga.H / 2) {
    if (x < 0) x += ga.W;
    if (y < 0) y += ga.H;
    p.onGrid = 1;
    p.x = x;
    p.y = y;
    p.kx = kx;
    p.ky = ky;
  }
  return p;
}

// Adds the number of lanes of the wave with `pred` to a workgroup counter in
// LDS (one atomic per wave; the workgroup flushes once to memory).
__device__ __forceinline__ void waveCount(unsigned long long* ldsCounter, bool pred) {
  const unsigned long long mask = __ballot(pred);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(ldsCounter, (unsigned long long)__popcll(mask));
}

// Consecutive lanes of a wave whose samples fall in the same bin form runs
// (samples are visited channel-fastest, so the channels of one row that stay
// in one tile are neighbours). The first lane of a run carries its length,
// so one atomic per run replaces one per sample: same-address atomics on the
// dense uv centre otherwise serialise in L2.
struct LaneRun {
  int head;  // lane index of this lane's run head
  int len;   // run length (valid on the head lane)
};

__device__ __forceinline__ LaneRun laneRun(uint32_t bin, bool valid) {
  const int lane = threadIdx.x & 63;
  const uint32_t prevBin = __shfl_up(bin, 1);
  const unsigned long long validMask = __ballot(valid);
  const bool prevValid = lane > 0 && ((validMask >> (lane - 1)) & 1ull);
  const bool head = valid && (lane == 0 || !prevValid || prevBin != bin);
  const unsigned long long heads = __ballot(head);
  const unsigned long long ends = heads | ~validMask;
  const unsigned long long after = (lane == 63) ? 0ull : ((ends >> (lane + 1)) << (lane + 1));
  const int next = after ? (__ffsll((long long)after) - 1) : 64;
  const unsigned long long upto = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
  LaneRun r;
  r.len = next - lane;
  r.head = upto ? (63 - __clzll((long long)upto)) : lane;
  return r;
}

struct KindArgs {
  LayerMap lm;
  uint32_t* counts;              // [nbins + 1]
  uint32_t* cursor;              // scatter: next free position per bin
  uint2* recSorted;              // scatter: records ordered by bin
  unsigned long long* counters;  // gridded, off grid, off layer
  uint32_t* sortKeys;            // scatter: bin << 12 | footprint origin, or null
  uint32_t nRecords;             // scatter: records of the plan
  uint32_t* skipped;             // scatter: samples no record grids [nSkipped]
  uint32_t nSkipped;
  unsigned long long* skippedCursor;
};

// The uv tiles a sample's footprint touches. Footprint cells run from
// (xs, ys) (centre minus K/2, wrapped into the grid) for K cells per axis;
// with tileSize >= K - 1 and tiles dividing the grid they touch at most two
// tiles per axis: tile A holding the first cell and, if different, tile B
// holding the last (possibly wrapped round to tile 0). For each tile the
// footprint start relative to the tile origin is xo in [-(K-1), T).
struct TileSpan {
  int tA, tB;    // tB < 0: single tile
  int xoA, xoB;
};

__device__ __forceinline__ TileSpan tileSpan(int xs, int K, int T, int tshift, int W) {
  TileSpan s;
  s.tA = tshift >= 0 ? (xs >> tshift) : xs / T;
  s.xoA = xs - s.tA * T;
  const int last = xs + K - 1;
  const int lastW = last >= W ? last - W : last;
  const int tl = tshift >= 0 ? (lastW >> tshift) : lastW / T;
  if (tl == s.tA && last < W) {
    s.tB = -1;
    s.xoB = 0;
  } else {
    s.tB = tl;
    s.xoB = (last < W) ? xs - tl * T : xs - W - tl * T;
  }
  return s;
}

// The tile spans of a placed sample's footprint along x and y.
__device__ __forceinline__ void footprintSpans(const GeomArgs& ga, const Placement& p,
                                               TileSpan& sx, TileSpan& sy) {
  int xs = p.x - ga.mid, ys = p.y - ga.mid;
  if (xs < 0) xs += ga.W;
  if (ys < 0) ys += ga.H;
  sx = tileSpan(xs, ga.K, ga.tileSize, ga.tileShift, ga.W);
  sy = tileSpan(ys, ga.K, ga.tileSize, ga.tileShift, ga.H);
}

// Bin and record of emission e (0: (A, A), 1: (B, A), 2: (A, B), 3: (B, B))
// of a placed sample with footprint spans sx, sy; returns false if that tile
// combination does not exist.
__device__ __forceinline__ bool emission(const GeomArgs& ga, const Placement& p,
                                         const TileSpan& sx, const TileSpan& sy, int e,
                                         uint32_t idx, uint32_t& bin, uint2& rec) {
  const bool useBx = e & 1, useBy = e & 2;
  if ((useBx && sx.tB < 0) || (useBy && sy.tB < 0)) return false;
  const int tx = useBx ? sx.tB : sx.tA, xo = useBx ? sx.xoB : sx.xoA;
  const int ty = useBy ? sy.tB : sy.tA, yo = useBy ? sy.xoB : sy.xoA;
  bin = uint32_t((size_t(p.layer) * ga.tilesY + ty) * ga.tilesX + tx);
  rec.x = idx | (uint32_t(e == 0) << 30) | (uint32_t(p.conj) << 31);
  rec.y = uint32_t(xo + ga.K - 1) | (uint32_t(yo + ga.K - 1) << 6) |
          (uint32_t(p.ky * ga.OS + p.kx) << 12);
  return true;
}

// Sort key of a row: the bin of the first tile its middle channel's footprint
// touches (0xFFFFFFFF when that sample is not gridded). Rows with equal keys
// have nearly equal uvw, so their samples fall in the same bins channel by
// channel.
__global__ void __launch_bounds__(kThreads)
rowKeyKernel(const double* __restrict__ uvw, const double* __restrict__ wl, size_t nrow,
             int nchan, GeomArgs ga, LayerMap lm, uint32_t* __restrict__ keys,
             uint32_t* __restrict__ rows) {
  const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  const double lambda = wl[nchan / 2];
  const Placement p = place(ga, lm, uvw[r * 3] / lambda, uvw[r * 3 + 1] / lambda,
                            uvw[r * 3 + 2] / lambda);
  uint32_t key = 0xFFFFFFFFu;
  if (p.layer >= 0 && p.onGrid) {
    TileSpan sx, sy;
    footprintSpans(ga, p, sx, sy);
    uint2 rec;
    emission(ga, p, sx, sy, 0, 0u, key, rec);
  }
  keys[r] = key;
  rows[r] = uint32_t(r);
}

constexpr int kBinSlots = 1024;      // LDS table of one group's bins
constexpr uint32_t kNoBin = 0xFFFFFFFFu;

// Slot of `bin` in a workgroup's LDS table, inserted if absent (linear
// probing from a multiplicative hash); -1 when 16 probes find neither the bin
// nor a free slot (the caller then counts in memory directly).
// The inserting thread appends the slot to `used` (count in *nUsed), so the
// table's flush and reset visit the used slots only.
__device__ __forceinline__ int binSlot(uint32_t* keys, uint32_t bin, uint16_t* used,
                                       uint32_t* nUsed) {
  uint32_t h = (bin * 2654435761u) >> 22;
  for (int probe = 0; probe < 16; ++probe) {
    const uint32_t k = *reinterpret_cast<volatile uint32_t*>(&keys[h]);
    if (k == bin) return int(h);
    if (k == kNoBin) {
      const uint32_t old = atomicCAS(&keys[h], kNoBin, bin);
      if (old == kNoBin) {
        used[atomicAdd(nUsed, 1u)] = uint16_t(h);
        return int(h);
      }
      if (old == bin) return int(h);
    }
    h = (h + 1u) & uint32_t(kBinSlots - 1);
  }
  return -1;
}

// Pass 1: place every sample (FP64, AddDataSample arithmetic) and histogram
// the (layer, uv tile) bins of all tiles its footprint touches. Pass 2
// (SCATTER) repeats the placement and writes the records to their bins.
// Samples are visited in row groups: group j holds positions
// [j * 1024, (j + 1) * 1024) of the samples in the order of `rowOrder` (rows
// sorted by their middle channel's bin, channels in order within a row), so
// the rows of a group share their bins. The lanes of a run of consecutive
// samples in one bin count it once; the runs of a group add their lengths in
// an LDS table, and the table adds one count per distinct bin to memory
// (the scatter: takes one block of positions per bin and hands each run its
// part). Workgroups visit groups in the order j = q * step mod ngroups
// (step coprime to ngroups), so groups in flight at once lie far apart in the
// sorted order and their bins rarely coincide. The gridded / off-grid /
// off-layer counters are summed per workgroup and added to memory once.
template <bool SCATTER, bool KEYS, int PER>
__global__ void __launch_bounds__(kThreads)
binKernel(const double* __restrict__ uvw, const double* __restrict__ wl,
          const uint32_t* __restrict__ rowOrder, size_t nrow, int nchan, uint32_t ngroups,
          uint32_t step, GeomArgs ga, KindArgs k) {
  __shared__ uint32_t slotBin[kBinSlots];
  __shared__ uint32_t slotCount[kBinSlots];  // scatter: then the bin's first position
  __shared__ uint16_t used[kBinSlots];       // the group's used slots
  __shared__ uint32_t nUsed[2];              // by group parity (reset one group later)
  __shared__ unsigned long long sCount[3];
  for (int t = threadIdx.x; t < kBinSlots; t += blockDim.x) {
    slotBin[t] = kNoBin;
    slotCount[t] = 0u;
  }
  if (threadIdx.x < 2) nUsed[threadIdx.x] = 0u;
  if (!SCATTER && threadIdx.x < 3) sCount[threadIdx.x] = 0ull;
  __syncthreads();
  const uint32_t n = uint32_t(nrow * size_t(nchan));
  const int lane = threadIdx.x & 63;
  uint32_t parity = 0u;
  for (uint32_t q = blockIdx.x; q < ngroups; q += gridDim.x, parity ^= 1u) {
    const uint32_t grp = uint32_t((uint64_t(q) * step) % ngroups);
    // per sample and emission: the record, and on the run head its table
    // slot and offset (or, past a full table, its position in memory)
    uint2 rec[PER][4];
    uint32_t meta[PER][4];  // bit 0 has, bit 1 direct, bits 2-7 head lane, 8-17 slot
    uint32_t loc[PER][4];   // head: offset within the slot, or the position
    uint32_t binKey[KEYS ? PER : 1][4];  // KEYS: the bin (for the sort keys)
#pragma unroll
    for (int qq = 0; qq < PER; ++qq) {
      const uint32_t s = grp * uint32_t(PER * kThreads) + uint32_t(qq * kThreads) + threadIdx.x;
      const bool inRange = s < n;
      uint32_t idx = 0u;
      Placement p;
      p.layer = -1;
      p.onGrid = 0;
      if (inRange) {
        const uint32_t sr = s / uint32_t(nchan);
        const uint32_t ch = s - sr * uint32_t(nchan);
        const uint32_t row = rowOrder[sr];
        idx = row * uint32_t(nchan) + ch;
        const double lambda = wl[ch];
        p = place(ga, k.lm, uvw[size_t(row) * 3] / lambda, uvw[size_t(row) * 3 + 1] / lambda,
                  uvw[size_t(row) * 3 + 2] / lambda);
      }
      const bool ok = inRange && p.layer >= 0 && p.onGrid;
      if (!SCATTER) {
        waveCount(&sCount[0], ok);
        waveCount(&sCount[1], inRange && p.layer >= 0 && !p.onGrid);
        waveCount(&sCount[2], inRange && p.layer < 0);
      } else {
        // list the samples no record grids (one atomic per wave)
        const bool skip = inRange && !ok;
        const unsigned long long sm = __ballot(skip);
        if (sm) {
          unsigned long long at0 = 0ull;
          if (lane == 0) at0 = atomicAdd(k.skippedCursor, (unsigned long long)__popcll(sm));
          at0 = __shfl(at0, 0);
          const unsigned long long at = at0 + __popcll(sm & ((1ull << lane) - 1ull));
          if (skip && at < k.nSkipped) k.skipped[at] = idx;
        }
      }
      TileSpan sx, sy;
      sx.tA = sx.tB = sy.tA = sy.tB = -1;
      sx.xoA = sx.xoB = sy.xoA = sy.xoB = 0;
      if (ok) footprintSpans(ga, p, sx, sy);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        uint32_t bin = kNoBin;
        uint2 r = make_uint2(0u, 0u);
        const bool has = ok && emission(ga, p, sx, sy, e, idx, bin, r);
        const LaneRun run = laneRun(bin, has);
        uint32_t m = has ? 1u : 0u;
        uint32_t l = 0u;
        if (has && run.head == lane) {
          const int slot = binSlot(slotBin, bin, used, &nUsed[parity]);
          if (slot >= 0) {
            l = atomicAdd(&slotCount[slot], uint32_t(run.len));
            m |= uint32_t(slot) << 8;
          } else {
            m |= 2u;
            if (!SCATTER)
              atomicAdd(&k.counts[bin], uint32_t(run.len));
            else
              l = atomicAdd(&k.cursor[bin], uint32_t(run.len));
          }
        }
        m |= uint32_t(run.head) << 2;
        rec[qq][e] = r;
        meta[qq][e] = m;
        loc[qq][e] = l;
        if constexpr (KEYS) binKey[qq][e] = bin;
      }
    }
    __syncthreads();
    // one count per distinct bin of the group (the used slots only)
    const uint32_t nu = nUsed[parity];
    for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) {
      const int t = used[u];
      const uint32_t bin = slotBin[t];
      if (!SCATTER) {
        atomicAdd(&k.counts[bin], slotCount[t]);
        slotBin[t] = kNoBin;
        slotCount[t] = 0u;
      } else {
        slotCount[t] = atomicAdd(&k.cursor[bin], slotCount[t]);
      }
    }
    // every thread has read nUsed[parity]; the next group counts in the
    // other one, and this one is cleared before the group after
    if (threadIdx.x == 0) nUsed[parity ^ 1u] = 0u;
    __syncthreads();
    if (SCATTER) {
#pragma unroll
      for (int qq = 0; qq < PER; ++qq)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t m = meta[qq][e];
          const int head = int((m >> 2) & 63u);
          uint32_t pos = 0u;
          if ((m & 1u) && head == lane)
            pos = (m & 2u) ? loc[qq][e] : slotCount[(m >> 8) & 1023u] + loc[qq][e];
          const uint32_t at = __shfl(pos, head) + uint32_t(lane - head);
          // (a position past the records is never written; the plan check
          // reports the bins it belongs to)
          if ((m & 1u) && at < k.nRecords) {
            k.recSorted[at] = rec[qq][e];
            if constexpr (KEYS) k.sortKeys[at] = (binKey[qq][e] << 12) | (rec[qq][e].y & 4095u);
          }
        }
      __syncthreads();
      for (uint32_t u = threadIdx.x; u < nu; u += blockDim.x) {
        const int t = used[u];
        slotBin[t] = kNoBin;
        slotCount[t] = 0u;
      }
      __syncthreads();
    }
  }
  if (!SCATTER) {
    __syncthreads();
    if (threadIdx.x < 3 && sCount[threadIdx.x])
      atomicAdd(&k.counters[threadIdx.x], sCount[threadIdx.x]);
  }
}

// Plan check: every bin's scatter cursor ends at the next bin's first record
// (each record was placed exactly once); counts the bins where it does not.
__global__ void __launch_bounds__(kThreads)
planCheckKernel(const uint32_t* __restrict__ cursor, const uint32_t* __restrict__ binStart,
                size_t nbins, unsigned long long* __restrict__ bad) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool wrong = b < nbins && cursor[b] != binStart[b + 1];
  const int cnt = __syncthreads_count(wrong);
  if (threadIdx.x == 0 && cnt) atomicAdd(bad, (unsigned long long)cnt);
}

// NaN for the images of one kind of every sample no record of that kind
// grids (the predict's "not predicted" value; the degridding writes all
// others).
__global__ void __launch_bounds__(kThreads)
nanSkippedKernel(const uint32_t* __restrict__ skipped, size_t n, ImgList imgs, int nImgTotal,
                 float2* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float q = __int_as_float(0x7FC00000);
  const size_t base = size_t(skipped[i]) * nImgTotal;
  for (int gi = 0; gi < imgs.n; ++gi) out[base + imgs.idx[gi]] = make_float2(q, q);
}

// Work items per bin: a bin of c > 0 records is processed as ceil(c / chunk)
// items, the first plus (c - 1) / chunk extra ones; an empty bin has none
// (its tile is cleared before gridding). Pass 1 writes the extra items per bin
// (and counts the non-empty bins), pass 2 -- after the scan of pass 1 into
// extraStart -- all items per bin.
__global__ void __launch_bounds__(kThreads)
extraItemsKernel(const uint32_t* __restrict__ binStart, size_t nbins, uint32_t chunk,
                 uint32_t* __restrict__ extra, unsigned long long* __restrict__ nonEmpty) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t c = 0;
  if (b < nbins) {
    c = binStart[b + 1] - binStart[b];
    extra[b] = c > chunk ? (c - 1u) / chunk : 0u;
  }
  const int cnt = __syncthreads_count(b < nbins && c > 0u);
  if (threadIdx.x == 0 && cnt) atomicAdd(nonEmpty, (unsigned long long)cnt);
}

__global__ void __launch_bounds__(kThreads)
binItemsKernel(const uint32_t* __restrict__ binStart, const uint32_t* __restrict__ extraStart,
               size_t nbins, uint32_t* __restrict__ items) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < nbins)
    items[b] = (binStart[b + 1] > binStart[b] ? 1u : 0u) + (extraStart[b + 1] - extraStart[b]);
}

// Item -> bin table: the bin of every work item (thread per bin; a split
// bin writes one entry per item).
__global__ void __launch_bounds__(kThreads)
itemBinsKernel(const uint32_t* __restrict__ itemStart, size_t nbins,
               uint32_t* __restrict__ itemBin) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  for (uint32_t j = itemStart[b], e = itemStart[b + 1]; j < e; ++j) itemBin[j] = uint32_t(b);
}

// Occupancy bits: bit b of the array is set when bin b holds records
// (thread per 32-bit word).
__global__ void __launch_bounds__(kThreads)
occupancyKernel(const uint32_t* __restrict__ binStart, size_t nbins,
                uint32_t* __restrict__ occ) {
  const size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t b0 = w * 32;
  if (b0 >= nbins) return;
  uint32_t bits = 0;
  for (int j = 0; j < 32 && b0 + j < nbins; ++j)
    if (binStart[b0 + j + 1] > binStart[b0 + j]) bits |= 1u << j;
  occ[w] = bits;
}

// Row bits: bit (layer * tilesY + ty) is set when some tile of that tile row
// holds records (thread per bit, packed by wave ballots).
__global__ void __launch_bounds__(kThreads)
rowLiveKernel(const uint32_t* __restrict__ occ, size_t nrows, int tilesX,
              uint32_t* __restrict__ rowLive) {
  const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool live = false;
  if (r < nrows) {
    const size_t b0 = r * size_t(tilesX);
    for (int t = 0; t < tilesX && !live; ++t) {
      const size_t b = b0 + t;
      live = (occ[b >> 5] >> (b & 31)) & 1u;
    }
  }
  const unsigned long long m = __ballot(live);
  const size_t w0 = (r & ~size_t(63)) >> 5;  // first word of this wave's 64 bits
  const int lane = threadIdx.x & 63;
  if (lane == 0 && w0 < (nrows + 31) / 32) rowLive[w0] = uint32_t(m);
  if (lane == 32 && w0 + 1 < (nrows + 31) / 32) rowLive[w0 + 1] = uint32_t(m >> 32);
}

// Adds the set bits of a bit array to *out (thread per 32-bit word).
__global__ void __launch_bounds__(kThreads)
popcountKernel(const uint32_t* __restrict__ bits, size_t words,
               unsigned long long* __restrict__ out) {
  const size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int c = w < words ? __popc(bits[w]) : 0;
  // wave sum, one atomic per wave
  int v = c;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, (unsigned long long)v);
}

// Column bits: bit (layer * tilesX + tx) is set when some tile row of that
// tile column has a set bit (thread per bit, packed by wave ballots): the
// column blocks a sparse forward columns pass reads, and the tile columns the
// forward rows pass writes.
__global__ void __launch_bounds__(kThreads)
liveColumnsKernel(const uint32_t* __restrict__ bits, size_t ncols, int tilesX, int tilesY,
                  uint32_t* __restrict__ colBits) {
  const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool live = false;
  if (c < ncols) {
    const size_t layer = c / size_t(tilesX), tx = c - layer * size_t(tilesX);
    for (int ty = 0; ty < tilesY && !live; ++ty) {
      const size_t b = (layer * size_t(tilesY) + size_t(ty)) * size_t(tilesX) + tx;
      live = (bits[b >> 5] >> (b & 31)) & 1u;
    }
  }
  const unsigned long long m = __ballot(live);
  const size_t w0 = (c & ~size_t(63)) >> 5;  // first word of this wave's 64 bits
  const int lane = threadIdx.x & 63;
  if (lane == 0 && w0 < (ncols + 31) / 32) colBits[w0] = uint32_t(m);
  if (lane == 32 && w0 + 1 < (ncols + 31) / 32) colBits[w0 + 1] = uint32_t(m >> 32);
}

// Need bits: tile (tx, ty) of a layer is read by the degridding subgrids of
// bins (tx, ty), (tx - 1, ty), (tx, ty - 1) and (tx - 1, ty - 1) (the
// (T + K - 1)^2 subgrid reaches K - 1 <= T cells into the next tiles,
// wrapping round the grid). Thread per 32-bit word.
__global__ void __launch_bounds__(kThreads)
needKernel(const uint32_t* __restrict__ occ, size_t nbins, int tilesX, int tilesY,
           uint32_t* __restrict__ need) {
  const size_t w = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t b0 = w * 32;
  if (b0 >= nbins) return;
  const size_t perLayer = size_t(tilesX) * tilesY;
  auto bit = [&](size_t b) { return (occ[b >> 5] >> (b & 31)) & 1u; };
  uint32_t bits = 0;
  for (int j = 0; j < 32 && b0 + j < nbins; ++j) {
    const size_t b = b0 + j;
    const size_t layerBase = (b / perLayer) * perLayer;
    const int t = int(b - layerBase);
    const int ty = t / tilesX, tx = t - ty * tilesX;
    const int px = tx == 0 ? tilesX - 1 : tx - 1, py = ty == 0 ? tilesY - 1 : ty - 1;
    const uint32_t v = bit(b) | bit(layerBase + size_t(ty) * tilesX + px) |
                       bit(layerBase + size_t(py) * tilesX + tx) |
                       bit(layerBase + size_t(py) * tilesX + px);
    bits |= v << j;
  }
  need[w] = bits;
}

// One work item of a pass: bin, item index within the bin (0 = the item
// that writes the tile) and its record range. The items of the pass's bins
// [binLo, binHi) are numbered in bin order from itemStart[binLo]; the bin of
// an item is found by binary search of the offsets.
struct WorkItem {
  uint32_t bin, j, start, count;
};

// Largest b in [lo, hi) with offs[b] <= v, for nondecreasing offs with
// offs[lo] <= v < offs[hi]: the bin that owns entry v (bins owning no entry
// are skipped because the next bin has the same offset).
__device__ __forceinline__ uint32_t ownerOf(const uint32_t* __restrict__ offs, uint32_t lo,
                                            uint32_t hi, uint32_t v) {
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ WorkItem decodeItem(uint32_t item, uint32_t binLo, uint32_t binHi,
                                               const uint32_t* __restrict__ binStart,
                                               const uint32_t* __restrict__ itemStart,
                                               uint32_t chunk,
                                               const uint32_t* __restrict__ itemBin = nullptr) {
  WorkItem w;
  const uint32_t g = itemStart[binLo] + item;
  w.bin = itemBin ? itemBin[g] : ownerOf(itemStart, binLo, binHi, g);
  w.j = g - itemStart[w.bin];
  const uint32_t s0 = binStart[w.bin], s1 = binStart[w.bin + 1];
  w.start = s0 + w.j * chunk;
  w.count = min(s1 - w.start, chunk);
  return w;
}

__device__ __forceinline__ uint32_t passItems(uint32_t binLo, uint32_t binHi,
                                              const uint32_t* __restrict__ itemStart) {
  return itemStart[binHi] - itemStart[binLo];
}


// --------------------------------------------------------- exclusive scan
constexpr int kScanItems = 8;
constexpr int kScanBlock = kThreads * kScanItems;  // 2048 values per block

__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v,
                                                       uint32_t* sh,
                                                       uint32_t* total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    uint32_t t = (tid >= off) ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  const uint32_t incl = sh[tid];
  *total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

// out[i] = sum_{j<i} in[j] within each block of kScanBlock values; block
// totals to blockSums.
__global__ void __launch_bounds__(kThreads)
scanBlocksKernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                 size_t n, uint32_t* __restrict__ blockSums) {
  __shared__ uint32_t sh[kThreads];
  const size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    vals[k] = (i < n) ? in[i] : 0u;
    local += vals[k];
  }
  uint32_t total;
  uint32_t run = blockExclusiveScan(local, sh, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    if (i < n) out[i] = run;
    run += vals[k];
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ void scanAddKernel(uint32_t* __restrict__ out, size_t n,
                              const uint32_t* __restrict__ blockOffsets) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += blockOffsets[i / kScanBlock];
}

// out[n] = total is written as well (out has n + 1 entries).
__global__ void scanTotalKernel(const uint32_t* __restrict__ in,
                                uint32_t* __restrict__ out, size_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    out[n] = (n > 0) ? out[n - 1] + in[n - 1] : 0u;
}

// ------------------------------------------------------------- grid (LDS)
struct GridArgs {
  GeomArgs ga;
  ImgList imgs;      // images of this kind
  int nImgTotal;     // images per sample in the visibility array
  int passStart;     // first layer of the pass
  int layerLo, layerHi;  // layers of the pass (absolute)
  int nLayersPass;   // allocated layers per image block
  int tilesPerLayer;
  int sg;            // degrid subgrid edge = tileSize + K - 1
  size_t nchan;
  uint32_t chunk;    // records per work item
  const float2* rot;  // rotating contexts: the gridding phasor of every sample
  // weighted contexts (hws_set_weighting): the weight every visibility is
  // gridded with, [sample][image] (visibilityWeight, formed by the plan's
  // weight-sum pass; null: the visibilities are gridded as given)
  const float* cw;
};

// The weight readAndWeightVisibilities multiplies a visibility by
// (gridding/msgridderbase.cpp:567-589, 640-659), formed once per plan by
// weightSumsKernel.
__device__ __forceinline__ float2 weighted(const GridArgs& a, float2 v, uint32_t idx, int img) {
  if (!a.cw) return v;
  const float w = a.cw[size_t(idx) * a.nImgTotal + img];
  return make_float2(v.x * w, v.y * w);
}

// The phase-centre rotation of a sample as a unit phasor, once per sample
// for all its images: MSGridderBase::rotateVisibilities
// (gridding/msgridderbase.cpp:692-704) with shiftFactor = sign * 2 pi
// (u l_shift + v m_shift) and rad = shiftFactor / lambda. The argument is
// FP64; the mixed precision reduces it to a fraction of a turn in FP64 and
// takes the FP32 turnSinCos, the reference precision an FP64 sincos.
__device__ __forceinline__ float2 rotationPhasor(const GridArgs& a,
                                                 const double* __restrict__ uvw,
                                                 const double* __restrict__ wl, uint32_t idx,
                                                 double sign) {
  const size_t row = idx / a.nchan, ch = idx - row * a.nchan;
  const double turns =
      sign * (uvw[row * 3] * a.ga.lShift + uvw[row * 3 + 1] * a.ga.mShift) / wl[ch];
  if (a.ga.fastPhase) {
    const float f = float(turns - rint(turns));
    float sf, cf;
    turnSinCos(f, sf, cf);
    return make_float2(cf, sf);
  }
  double s, c;
  sincos(2.0 * kPi * turns, &s, &c);
  return make_float2(float(c), float(s));
}

__device__ __forceinline__ float2 rotateBy(float2 v, float2 p) {
  return make_float2(v.x * p.x - v.y * p.y, v.x * p.y + v.y * p.x);
}

// The gridding phasor of every sample of a plan (rotationPhasor with sign
// -1), once per plan: the grid and degrid kernels read it per record
// instead of evaluating the FP64 argument and its sine and cosine for each
// of a sample's records and image kinds.
__global__ void __launch_bounds__(256)
rotationTableKernel(GridArgs a, const double* __restrict__ uvw, const double* __restrict__ wl,
                    size_t n, float2* __restrict__ rot) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) rot[i] = rotationPhasor(a, uvw, wl, uint32_t(i), -1.0);
}

// Kernel tables into LDS: FP32 tables, or FP16 (hi, lo) pairs whose sum
// restores the FP32 value to about 22 bits.
__device__ __forceinline__ void loadKernelTable(float* kt, const float* __restrict__ ktab,
                                                const __half2* __restrict__ khalf, int n) {
  if (khalf) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const __half2 h = khalf[i];
      kt[i] = __low2float(h) + __high2float(h);
    }
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) kt[i] = ktab[i];
  }
}

// Orders the LDS accesses of one wave before and after this point (lanes of
// a wave exchanging data through LDS).
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kGridWaves = 2;
constexpr int kGridThreads = 64 * kGridWaves;

// Row pitch of a wave's private tile: at least T, and congruent to +-7
// modulo 32, so that the 7 rows of a 7 x 7 footprint fall on different
// banks (at most two taps per bank).
__host__ __device__ inline int privatePitch(int T) {
  int p = T;
  while ((p & 31) != 7 && (p & 31) != 25) ++p;
  return p;
}

// Gridding of one image kind, one work item per workgroup at a time: the
// workgroup owns the T x T tile of the item's (layer, uv tile) bin. Each wave
// takes a contiguous segment of the item's records and walks them one at a
// time with lane t owning tap t of the K x K footprint (TG groups of 64 taps
// for K > 8); the records' positions and visibilities are staged in LDS 64
// at a time and read by every lane. While consecutive records keep the
// footprint in place (neighbouring channels of a short baseline) each lane
// sums in registers; when it moves, the lanes add their sums to the wave's
// own copy of the tile in LDS with plain read-modify-write (the taps of one
// footprint are distinct cells and no other wave writes the copy, so no
// atomics are needed; LDS float atomics run far below LDS bandwidth). The
// copies are then summed and the tile written once with plain stores:
// item 0 of a bin to the layer (ACCUM adds to what the layer holds, for
// callers that grid a pass in several batches), extra items of a large bin
// to their partial tile, which reducePartialsKernel adds to the layer
// afterwards in a fixed order. No global atomics: tiles without samples are
// written as zeros, which also clears the layer.
template <int NIMG, int TG, bool ACCUM>
__global__ void __launch_bounds__(kGridThreads)
gridTilesKernel(GridArgs a, const uint32_t* __restrict__ binStart,
                const uint32_t* __restrict__ extraStart, const uint32_t* __restrict__ itemStart,
                const uint32_t* __restrict__ itemBin, const uint2* __restrict__ recs,
                const float2* __restrict__ vis, const float* __restrict__ ktab,
                const __half2* __restrict__ khalf, const double* __restrict__ uvw,
                const double* __restrict__ wl, float2* __restrict__ layers,
                float2* __restrict__ partials) {
  extern __shared__ float smem[];
  const int K = a.ga.K, OS = a.ga.OS, T = a.ga.tileSize, T2 = T * T;
  const int P = privatePitch(T), TP = T * P;  // private tile: T rows of pitch P
  float* kt = smem;                            // [OS][K]
  float* priv = smem + ((OS * K + 3) & ~3);    // [wave][NIMG][re, im][T][P]
  // staged records of each wave: position key, kernel row offsets, values
  int2* stPos = reinterpret_cast<int2*>(priv + kGridWaves * NIMG * 2 * TP);  // [wave][64]
  float2* stVis = reinterpret_cast<float2*>(stPos + kGridWaves * 64);        // [wave][64][NIMG]
  loadKernelTable(kt, ktab, khalf, OS * K);
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* myTile = priv + wave * NIMG * 2 * TP;
  int2* myPos = stPos + wave * 64;
  float2* myVis = stVis + wave * 64 * NIMG;
  const int taps = K * K;
  int tj[TG], ti[TG];
  bool tapOk[TG];
#pragma unroll
  for (int q = 0; q < TG; ++q) {
    const int t = lane + 64 * q;
    tapOk[q] = t < taps;
    tj[q] = tapOk[q] ? t / K : 0;
    ti[q] = tapOk[q] ? t - tj[q] * K : 0;
  }
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  const uint32_t nItems = passItems(binLo, binHi, itemStart);
  __syncthreads();
  for (uint32_t item = blockIdx.x; item < nItems; item += gridDim.x) {
    // every lane decodes the item (uniform loads)
    const WorkItem wi = decodeItem(item, binLo, binHi, binStart, itemStart, a.chunk, itemBin);
    const uint32_t bin = wi.bin, start = wi.start, count = wi.count;
    if (ACCUM && count == 0) continue;  // block-uniform; nothing to add
    const int layer = int(bin / uint32_t(a.tilesPerLayer));
    const int rem = int(bin - uint32_t(layer) * uint32_t(a.tilesPerLayer));
    const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
    if (count > 0) {
      // the wave's own copy: no other wave reads or writes it until the sum
      for (int i = lane; i < NIMG * 2 * TP; i += 64) myTile[i] = 0.0f;
      waveSync();
      // wave segment of the item's records
      const uint32_t seg = (count + kGridWaves - 1) / kGridWaves;
      const uint32_t w0 = min(count, uint32_t(wave) * seg), w1 = min(count, w0 + seg);
      float acc[TG][2 * NIMG];
#pragma unroll
      for (int q = 0; q < TG; ++q)
#pragma unroll
        for (int e = 0; e < 2 * NIMG; ++e) acc[q][e] = 0.0f;
      int cur = -1;
      for (uint32_t b0 = w0; b0 < w1; b0 += 64u) {
        // stage up to 64 records: lane l loads record b0 + l
        const uint32_t it = b0 + uint32_t(lane);
        if (it < w1) {
          const uint2 rec = recs[start + it];
          const uint32_t idx = rec.x & 0x3FFFFFFFu;
          const bool conj = (rec.x >> 31) != 0;
          const int kk = int(rec.y >> 12);
          const int ky = kk / OS;
          myPos[lane] = make_int2(int(rec.y & 4095u), (ky * K) | ((kk - ky * OS) * K) << 16);
          const float2 ph = a.ga.rotate ? a.rot[idx] : make_float2(1.0f, 0.0f);
#pragma unroll
          for (int gi = 0; gi < NIMG; ++gi) {
            float2 v = weighted(a, vis[size_t(idx) * a.nImgTotal + a.imgs.idx[gi]], idx,
                                a.imgs.idx[gi]);
            if (a.ga.rotate) v = rotateBy(v, ph);
            if (conj) v.y = -v.y;
            myVis[lane * NIMG + gi] = v;
          }
        }
        waveSync();
        const int nb = int(min(64u, w1 - b0));
        // eight records at a time: their staged data and kernel values are
        // read from LDS together, so one LDS latency covers eight records
        for (int s0 = 0; s0 < nb; s0 += 8) {
          int key[8];
          float kv[8][TG];
          float2 v[8][NIMG];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int su = min(s0 + u, nb - 1);
            const int2 ps = myPos[su];  // same address in every lane: broadcast
            key[u] = (s0 + u < nb) ? ps.x : -2;
            const int kyo = ps.y & 0xFFFF, kxo = ps.y >> 16;
#pragma unroll
            for (int q = 0; q < TG; ++q)
              // yKernelValue * xKernel[i] (gridding/wstackinggridder.cpp:574)
              kv[u][q] = kt[kyo + tj[q]] * kt[kxo + ti[q]];
#pragma unroll
            for (int gi = 0; gi < NIMG; ++gi) v[u][gi] = myVis[su * NIMG + gi];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (key[u] == -2) break;  // past the batch end (wave-uniform)
            if (key[u] != cur) {
              if (cur >= 0) {
                const int xo = (cur & 63) - (K - 1), yo = (cur >> 6) - (K - 1);
#pragma unroll
                for (int q = 0; q < TG; ++q) {
                  const int cx = xo + ti[q], cy = yo + tj[q];
                  if (tapOk[q] && cx >= 0 && cx < T && cy >= 0 && cy < T) {
                    float* cell = myTile + cy * P + cx;
                    float old[2 * NIMG];  // all loads first: one LDS latency
#pragma unroll
                    for (int e = 0; e < 2 * NIMG; ++e) old[e] = cell[e * TP];
#pragma unroll
                    for (int e = 0; e < 2 * NIMG; ++e) cell[e * TP] = old[e] + acc[q][e];
                  }
#pragma unroll
                  for (int e = 0; e < 2 * NIMG; ++e) acc[q][e] = 0.0f;
                }
              }
              cur = key[u];
            }
#pragma unroll
            for (int q = 0; q < TG; ++q)
#pragma unroll
              for (int gi = 0; gi < NIMG; ++gi) {
                acc[q][2 * gi] += v[u][gi].x * kv[u][q];
                acc[q][2 * gi + 1] += v[u][gi].y * kv[u][q];
              }
          }
        }
        waveSync();  // staging is rewritten next batch
      }
      if (cur >= 0) {
        const int xo = (cur & 63) - (K - 1), yo = (cur >> 6) - (K - 1);
#pragma unroll
        for (int q = 0; q < TG; ++q) {
          const int cx = xo + ti[q], cy = yo + tj[q];
          if (tapOk[q] && cx >= 0 && cx < T && cy >= 0 && cy < T) {
            float* cell = myTile + cy * P + cx;
            float old[2 * NIMG];
#pragma unroll
            for (int e = 0; e < 2 * NIMG; ++e) old[e] = cell[e * TP];
#pragma unroll
            for (int e = 0; e < 2 * NIMG; ++e) cell[e * TP] = old[e] + acc[q][e];
          }
        }
      }
    }
    __syncthreads();  // every wave's copy is complete
    // sum of the copies, cell c of image gi (0 for an empty bin)
    auto tileValue = [&](int gi, int cx, int cy) {
      float2 v = make_float2(0.0f, 0.0f);
      if (count > 0) {
        const int o = cy * P + cx;
#pragma unroll
        for (int w = 0; w < kGridWaves; ++w) {
          const float* t = priv + w * NIMG * 2 * TP + (2 * gi) * TP;
          v.x += t[o];
          v.y += t[TP + o];
        }
      }
      return v;
    };
    if (wi.j > 0) {
      // an extra item of a large bin: its tile goes to the partial slot
      float2* dst = partials + size_t(extraStart[bin] + wi.j - 1u) * size_t(NIMG * T2);
      for (int c = threadIdx.x; c < T2; c += blockDim.x) {
        const int cy = c / T, cx = c - cy * T;
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) dst[gi * T2 + c] = tileValue(gi, cx, cy);
      }
    } else {
      // write the tile in layer storage order (blocked: 4 rows x T columns
      // contiguous per row block)
      const int x0 = tx * T, y0 = ty * T;
      for (int c = threadIdx.x; c < T2; c += blockDim.x) {
        int cx, cy;
        if (a.ga.blocked) {
          const int rb = c / (4 * T), r4 = c - rb * 4 * T;
          cx = r4 >> 2;
          cy = rb * 4 + (r4 & 3);
        } else {
          cy = c / T;
          cx = c - cy * T;
        }
        const size_t off = cellIndex(x0 + cx, y0 + cy, a.ga.W, a.ga.blocked);
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          float2* dst = layers +
                        (size_t(a.imgs.idx[gi]) * a.nLayersPass + (layer - a.passStart)) * HW + off;
          float2 v = tileValue(gi, cx, cy);
          if (ACCUM) {
            const float2 o = *dst;
            v.x += o.x;
            v.y += o.y;
          }
          *dst = v;
        }
      }
    }
    __syncthreads();  // the copies are read before the next item clears them
  }
}

constexpr int kCellWaves = 4;
constexpr int kCellThreads = 64 * kCellWaves;

// Tile edges the cell-owning grid kernel takes (64 lanes cover whole rows).
__host__ __device__ inline bool cellKernelTile(int T) { return T == 8 || T == 16; }

// Row length of the zero-padded kernel table of the cell kernel: K values
// with T - 1 zeros on each side, so that any tap offset a cell of the tile
// can have from a footprint start reads either a kernel value or a zero.
__host__ __device__ inline int paddedRow(int K, int T) { return K + 2 * (T - 1); }

// The padded table is used when it fits in 32 KiB of LDS (OS = 63 at K = 7);
// larger oversampling uses the compact table with clamped indexes.
__host__ __device__ inline bool cellTablePadded(int K, int OS, int T) {
  return size_t(OS) * paddedRow(K, T) * sizeof(float) <= 32768;
}

// Gridding of one image kind, one work item per WAVE: the wave owns the T x T
// tile of the item's (layer, uv tile) bin with lane l holding the cells of
// column l mod T and R consecutive rows from (l div T) * R, so 64 lanes cover
// the tile (T = 16: 4 rows per lane; T = 8: 1). Records are loaded 64 at a
// time, one per lane, into registers (kernel-table offsets, rotated and
// conjugated visibilities) and broadcast to the wave one by one through
// scalar registers; every lane then adds to its own cells the taps of the
// record's K x K footprint that fall on them, kernel value
// yKernel[dy] * xKernel[dx] (gridding/wstackinggridder.cpp:560-597). With
// the zero-padded table (PAD) a cell outside the footprint reads a zero, so
// the loop has no tests at all. ROT applies the phase-centre rotation (a
// separate instance: its FP64 sincos costs registers). The sums stay in
// registers for the whole
// item: no LDS traffic but the kernel table, no read-modify-write chains, no
// atomics. Item 0 of a bin writes the tile to the layer (ACCUM adds to it);
// extra items of a split bin write partial tiles that reducePartialsKernel
// adds in item order.
template <int NIMG, int T, bool ACCUM, bool PAD, bool ROT>
__global__ void __launch_bounds__(kCellThreads)
gridCellsKernel(GridArgs a, const uint32_t* __restrict__ binStart,
                const uint32_t* __restrict__ extraStart, const uint32_t* __restrict__ itemStart,
                const uint32_t* __restrict__ itemBin, const uint2* __restrict__ recs,
                const float2* __restrict__ vis, const float* __restrict__ ktab,
                const __half2* __restrict__ khalf, const double* __restrict__ uvw,
                const double* __restrict__ wl, float2* __restrict__ layers,
                float2* __restrict__ partials, uint32_t* __restrict__ queue) {
  constexpr int G = 64 / T;  // lane row groups
  constexpr int R = T / G;   // rows per lane
  static_assert(G * T == 64 && G * R == T, "64 lanes must cover the tile");
  constexpr int T2 = T * T;
  extern __shared__ float smem[];
  const int K = a.ga.K, OS = a.ga.OS;
  float* kt = smem;  // PAD: [OS][K + 2 (T - 1)], else [OS][K]
  const int rowLen = PAD ? paddedRow(K, T) : K;
  if (PAD) {
    for (int i = threadIdx.x; i < OS * rowLen; i += blockDim.x) {
      const int row = i / rowLen, c = i - row * rowLen - (T - 1);
      float v = 0.0f;
      if (c >= 0 && c < K) {
        const int src = row * K + c;
        if (khalf) {
          const __half2 h = khalf[src];
          v = __low2float(h) + __high2float(h);
        } else {
          v = ktab[src];
        }
      }
      kt[i] = v;
    }
  } else {
    loadKernelTable(kt, ktab, khalf, OS * K);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cx = lane % T, cy0 = (lane / T) * R;
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const uint32_t nItems = passItems(binLo, binHi, itemStart);
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  // items come from a work queue (a counter the host clears before the
  // launch): a wave takes the next item when it finishes one, so a few
  // heavy items do not leave most waves idle at the end. The next index is
  // requested while the current item runs.
  (void)wave;
  auto take = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(queue, 1u);
    return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
  };
  uint32_t next = take();
  while (next < nItems) {
    const uint32_t item = next;
    const WorkItem wi = decodeItem(item, binLo, binHi, binStart, itemStart, a.chunk, itemBin);
    next = take();
    const uint32_t bin = wi.bin, start = wi.start, count = wi.count;
    if (ACCUM && count == 0) continue;
    float acc[R][2 * NIMG];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int e = 0; e < 2 * NIMG; ++e) acc[r][e] = 0.0f;
    for (uint32_t b0 = 0; b0 < count; b0 += 64u) {
      // lane l loads record b0 + l; lanes past the end hold a zero sample
      // (PAD: table offsets of a zero row region; else position and rows)
      int pos = 0, kof = 0;
      float2 v[NIMG];
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) v[gi] = make_float2(0.0f, 0.0f);
      const uint32_t it = b0 + uint32_t(lane);
      if (it < count) {
        const uint2 rec = recs[start + it];
        const uint32_t idx = rec.x & 0x3FFFFFFFu;
        const bool conj = (rec.x >> 31) != 0;
        const int kk = int(rec.y >> 12);
        const int ky = kk / OS, kx = kk - ky * OS;
        const int xs = int(rec.y & 63u), ys = int((rec.y >> 6) & 63u);  // origin + K - 1
        if (PAD) {
          // table index of tap (cx, cy) = row base + (T - 1) + (c - origin)
          pos = (kx * rowLen + (T - 1) + (K - 1) - xs) |
                ((ky * rowLen + (T - 1) + (K - 1) - ys) << 16);
        } else {
          pos = xs | (ys << 6);
          kof = (ky * K) | (kx * K) << 16;
        }
        float2 ph = make_float2(1.0f, 0.0f);
        if (ROT) ph = a.rot[idx];
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          float2 x = weighted(a, vis[size_t(idx) * a.nImgTotal + a.imgs.idx[gi]], idx,
                              a.imgs.idx[gi]);
          if (ROT) x = rotateBy(x, ph);
          if (conj) x.y = -x.y;
          v[gi] = x;
        }
      }
      const int nb = int(min(64u, count - b0));
      for (int u0 = 0; u0 < nb; u0 += 4) {
        // four records at a time: every kernel-table read is unconditional,
        // so the reads of all four issue before the first wait
        float kx[4], ky[4][R];
        bool inx[4], iny[4][R];
#pragma unroll
        for (int du = 0; du < 4; ++du) {
          const int u = u0 + du;
          const int ps = __builtin_amdgcn_readlane(pos, u);
          if (PAD) {
            const int ax = (ps & 0xFFFF) + cx, ay = (ps >> 16) + cy0;
            kx[du] = kt[ax];
#pragma unroll
            for (int r = 0; r < R; ++r) ky[du][r] = kt[ay + r];
          } else {
            const int ko = __builtin_amdgcn_readlane(kof, u);
            const int xo = (ps & 63) - (K - 1), yo = (ps >> 6) - (K - 1);
            const int kyo = ko & 0xFFFF, kxo = ko >> 16;
            const int dx = cx - xo;
            inx[du] = unsigned(dx) < unsigned(K);
            kx[du] = kt[kxo + min(max(dx, 0), K - 1)];
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int dy = cy0 + r - yo;
              iny[du][r] = unsigned(dy) < unsigned(K);
              ky[du][r] = kt[kyo + min(max(dy, 0), K - 1)];
            }
          }
        }
#pragma unroll
        for (int du = 0; du < 4; ++du) {
          const int u = u0 + du;
          float2 vv[NIMG];
#pragma unroll
          for (int gi = 0; gi < NIMG; ++gi) {
            vv[gi].x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[gi].x), u));
            vv[gi].y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[gi].y), u));
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            // yKernelValue * xKernel[i] (gridding/wstackinggridder.cpp:574)
            float kv = ky[du][r] * kx[du];
            if (!PAD) kv = (inx[du] && iny[du][r]) ? kv : 0.0f;
#pragma unroll
            for (int gi = 0; gi < NIMG; ++gi) {
              acc[r][2 * gi] += vv[gi].x * kv;
              acc[r][2 * gi + 1] += vv[gi].y * kv;
            }
          }
        }
      }
    }
    if (wi.j > 0) {
      // an extra item of a large bin: its tile goes to the partial slot
      float2* dst = partials + size_t(extraStart[bin] + wi.j - 1u) * size_t(NIMG * T2);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi)
          dst[gi * T2 + (cy0 + r) * T + cx] = make_float2(acc[r][2 * gi], acc[r][2 * gi + 1]);
    } else {
      const int layer = int(bin / uint32_t(a.tilesPerLayer));
      const int rem = int(bin - uint32_t(layer) * uint32_t(a.tilesPerLayer));
      const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
      const int x = tx * T + cx, y0 = ty * T + cy0;
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) {
        float2* base = layers + (size_t(a.imgs.idx[gi]) * a.nLayersPass +
                                 size_t(layer - a.passStart)) * HW;
        if (R == 4 && a.ga.blocked) {
          // rows y0 .. y0 + 3 of one column are 32 contiguous bytes
          float4* dst = reinterpret_cast<float4*>(base + cellIndex(x, y0, a.ga.W, 1));
          float4 lo = make_float4(acc[0][2 * gi], acc[0][2 * gi + 1], acc[1 % R][2 * gi],
                                  acc[1 % R][2 * gi + 1]);
          float4 hi = make_float4(acc[2 % R][2 * gi], acc[2 % R][2 * gi + 1],
                                  acc[3 % R][2 * gi], acc[3 % R][2 * gi + 1]);
          if (ACCUM) {
            const float4 o0 = dst[0], o1 = dst[1];
            lo.x += o0.x; lo.y += o0.y; lo.z += o0.z; lo.w += o0.w;
            hi.x += o1.x; hi.y += o1.y; hi.z += o1.z; hi.w += o1.w;
          }
          dst[0] = lo;
          dst[1] = hi;
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            float2* dst = base + cellIndex(x, y0 + r, a.ga.W, a.ga.blocked);
            float2 val = make_float2(acc[r][2 * gi], acc[r][2 * gi + 1]);
            if (ACCUM) {
              const float2 o = *dst;
              val.x += o.x;
              val.y += o.y;
            }
            *dst = val;
          }
        }
      }
    }
  }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// The tile of a bin as a matrix product: with A[y][r] = v_r * yKernel_r[y -
// y0_r] and B[r][x] = xKernel_r[x - x0_r] (zero outside record r's
// footprint), the gridded 16 x 16 tile is G = sum_r A[.][r] B[r][.] = A B,
// a rank-K update over the item's records, done four records at a time by
// v_mfma_f32_16x16x4_f32 (FP32 products and sums) per image component.
// Operand layout of that instruction: lane i + 16 k supplies A[i][k] and
// B[k][i]; lane j + 16 (i / 4) receives D[i][j] in register i % 4, which is
// the cell ownership of gridCellsKernel at T = 16 (column lane mod 16, rows
// 4 (lane div 16) ..+3). Records are staged in the wave's LDS slice 64 at a
// time; lane i + 16 k reads record u0 + k of each group of four and looks
// up its own row (ky) and column (kx) in the zero-padded kernel table.
// Work items come from the work queue as in gridCellsKernel.
constexpr int kMfmaWaves = 4;
constexpr int kMfmaThreads = 64 * kMfmaWaves;

template <int NIMG, bool ACCUM, bool ROT>
__global__ void __launch_bounds__(kMfmaThreads)
gridMfmaKernel(GridArgs a, const uint32_t* __restrict__ binStart,
               const uint32_t* __restrict__ extraStart, const uint32_t* __restrict__ itemStart,
               const uint32_t* __restrict__ itemBin, const uint2* __restrict__ recs,
               const float2* __restrict__ vis, const float* __restrict__ ktab,
               const __half2* __restrict__ khalf, const double* __restrict__ uvw,
               const double* __restrict__ wl, float2* __restrict__ layers,
               float2* __restrict__ partials, uint32_t* __restrict__ queue) {
  constexpr int T = 16, T2 = T * T;
  extern __shared__ float smem[];
  const int K = a.ga.K, OS = a.ga.OS;
  const int rowLen = paddedRow(K, T);
  float* kt = smem;  // [OS][K + 2 (T - 1)], zero padded
  const int tableFloats = (OS * rowLen + 3) & ~3;
  int* stPos = reinterpret_cast<int*>(smem + tableFloats);                 // [wave][64]
  float2* stVis = reinterpret_cast<float2*>(stPos + kMfmaWaves * 64);      // [wave][64][NIMG]
  for (int i = threadIdx.x; i < OS * rowLen; i += blockDim.x) {
    const int row = i / rowLen, c = i - row * rowLen - (T - 1);
    float v = 0.0f;
    if (c >= 0 && c < K) {
      const int src = row * K + c;
      if (khalf) {
        const __half2 h = khalf[src];
        v = __low2float(h) + __high2float(h);
      } else {
        v = ktab[src];
      }
    }
    kt[i] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;  // operand row/column and record slot
  int* myPos = stPos + wave * 64;
  float2* myVis = stVis + wave * 64 * NIMG;
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const uint32_t nItems = passItems(binLo, binHi, itemStart);
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  auto take = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(queue, 1u);
    return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
  };
  uint32_t next = take();
  while (next < nItems) {
    const uint32_t item = next;
    const WorkItem wi = decodeItem(item, binLo, binHi, binStart, itemStart, a.chunk, itemBin);
    next = take();
    const uint32_t bin = wi.bin, start = wi.start, count = wi.count;
    if (ACCUM && count == 0) continue;
    floatx4 acc[2 * NIMG];
#pragma unroll
    for (int e = 0; e < 2 * NIMG; ++e) acc[e] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t b0 = 0; b0 < count; b0 += 64u) {
      // stage 64 records (zero samples past the end: their A rows vanish)
      int pos = 0;
      float2 v[NIMG];
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) v[gi] = make_float2(0.0f, 0.0f);
      const uint32_t it = b0 + uint32_t(lane);
      if (it < count) {
        const uint2 rec = recs[start + it];
        const uint32_t idx = rec.x & 0x3FFFFFFFu;
        const bool conj = (rec.x >> 31) != 0;
        const int kk = int(rec.y >> 12);
        const int ky = kk / OS, kx = kk - ky * OS;
        const int xs = int(rec.y & 63u), ys = int((rec.y >> 6) & 63u);
        pos = (kx * rowLen + (T - 1) + (K - 1) - xs) |
              ((ky * rowLen + (T - 1) + (K - 1) - ys) << 16);
        float2 ph = make_float2(1.0f, 0.0f);
        if (ROT) ph = a.rot[idx];
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          float2 x = weighted(a, vis[size_t(idx) * a.nImgTotal + a.imgs.idx[gi]], idx,
                              a.imgs.idx[gi]);
          if (ROT) x = rotateBy(x, ph);
          if (conj) x.y = -x.y;
          v[gi] = x;
        }
      }
      myPos[lane] = pos;
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) myVis[lane * NIMG + gi] = v[gi];
      waveSync();
      // all 64 staged records, sixteen (four groups of four) per round with
      // their LDS reads issued before the first product; the staging is zero
      // past the item's end, and a fully unrolled batch keeps the
      // accumulators in place (a partial last batch costs at most 63 zero
      // records)
#pragma unroll
      for (int u0 = 0; u0 < 64; u0 += 16) {
        int ps[4];
        float2 vv[4][NIMG];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = u0 + 4 * q + lk;
          ps[q] = myPos[r];
#pragma unroll
          for (int gi = 0; gi < NIMG; ++gi) vv[q][gi] = myVis[r * NIMG + gi];
        }
        float bx[4], ky[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          // B[k][x = li] = xKernel, A[y = li][k] = v * yKernel
          // (yKernelValue * xKernel[i], gridding/wstackinggridder.cpp:574)
          bx[q] = kt[(ps[q] & 0xFFFF) + li];
          ky[q] = kt[(ps[q] >> 16) + li];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int gi = 0; gi < NIMG; ++gi) {
            acc[2 * gi] = __builtin_amdgcn_mfma_f32_16x16x4f32(vv[q][gi].x * ky[q], bx[q],
                                                               acc[2 * gi], 0, 0, 0);
            acc[2 * gi + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(vv[q][gi].y * ky[q], bx[q],
                                                                   acc[2 * gi + 1], 0, 0, 0);
          }
      }
      waveSync();  // the staging is rewritten next batch
    }
    // lane owns column li, rows 4 lk .. 4 lk + 3
    if (wi.j > 0) {
      float2* dst = partials + size_t(extraStart[bin] + wi.j - 1u) * size_t(NIMG * T2);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi)
          dst[gi * T2 + (4 * lk + r) * T + li] = make_float2(acc[2 * gi][r], acc[2 * gi + 1][r]);
    } else {
      const int layer = int(bin / uint32_t(a.tilesPerLayer));
      const int rem = int(bin - uint32_t(layer) * uint32_t(a.tilesPerLayer));
      const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
      const int x = tx * T + li, y0 = ty * T + 4 * lk;
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) {
        float2* base = layers + (size_t(a.imgs.idx[gi]) * a.nLayersPass +
                                 size_t(layer - a.passStart)) * HW;
        const floatx4 re = acc[2 * gi], im = acc[2 * gi + 1];
        if (a.ga.blocked) {
          // rows y0 .. y0 + 3 of one column are 32 contiguous bytes
          float4* dst = reinterpret_cast<float4*>(base + cellIndex(x, y0, a.ga.W, 1));
          float4 lo = make_float4(re[0], im[0], re[1], im[1]);
          float4 hi = make_float4(re[2], im[2], re[3], im[3]);
          if (ACCUM) {
            const float4 o0 = dst[0], o1 = dst[1];
            lo.x += o0.x; lo.y += o0.y; lo.z += o0.z; lo.w += o0.w;
            hi.x += o1.x; hi.y += o1.y; hi.z += o1.z; hi.w += o1.w;
          }
          dst[0] = lo;
          dst[1] = hi;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float2* dst = base + cellIndex(x, y0 + r, a.ga.W, 0);
            float2 val = make_float2(re[r], im[r]);
            if (ACCUM) {
              const float2 o = *dst;
              val.x += o.x;
              val.y += o.y;
            }
            *dst = val;
          }
        }
      }
    }
  }
}

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// Ranges of the split operands x = h + l (h the FP16 rounding of x, l that
// of x - h): |A| < 2^14, |B| < 2^13, so that both parts stay below FP16's
// maximum (65504), and a low part lost to FP16's subnormal range (|l| <
// 2^-14, |x| below about 2^-3) is below 2^-27 of the largest operand.
constexpr int kAExp = 14, kBExp = 13;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Entry i of the FP16 grid kernel's padded table (rows of paddedRow(K, 16)
// values, the kernel row at offset T - 1, zeros around it): the FP32 value,
// and the (h, l) FP16 split of it times bScale packed in one word.
__device__ __forceinline__ void f16TableEntry(int i, const float* __restrict__ ktab, int K,
                                              int rowLen, float bScale, float& v,
                                              uint32_t& pair) {
#pragma clang fp contract(off)
  constexpr int T = 16;
  const int row = i / rowLen, c = i - row * rowLen - (T - 1);
  v = 0.0f;
  if (c >= 0 && c < K) v = ktab[row * K + c];
  const float vs = v * bScale;
  const _Float16 h = (_Float16)vs;
  const _Float16 l = (_Float16)(vs - (float)h);
  pair = uint32_t(__builtin_bit_cast(uint16_t, h)) |
         (uint32_t(__builtin_bit_cast(uint16_t, l)) << 16);
}

// The padded table in device memory, for oversampling factors whose table
// does not fit the kernel's LDS (read through the caches instead).
__global__ void __launch_bounds__(kThreads)
f16TableKernel(const float* __restrict__ ktab, int K, int OS, int eB, float* __restrict__ kt,
               uint32_t* __restrict__ kp) {
  const int rowLen = paddedRow(K, 16);
  const int i = int(blockIdx.x) * kThreads + int(threadIdx.x);
  if (i >= OS * rowLen) return;
  f16TableEntry(i, ktab, K, rowLen, ldexpf(1.0f, kBExp - eB), kt[i], kp[i]);
}

// The FP16 roundings of x0 - h0 and x1 - h1 packed into one word, h0 and h1
// the two halves of h (x - h is exact in FP32 where h is x rounded toward
// zero; v_fma_mix forms it from the FP16 operand and rounds once). Written
// by hand: the compiler does not select v_fma_mix for it.
__device__ __forceinline__ uint32_t lowParts(float x0, float x1, uint32_t h) {
  uint32_t lo;
  __asm__("v_fma_mixlo_f16 %0, %1, 1.0, -%3 op_sel_hi:[0,0,1]\n\t"
          "v_fma_mixhi_f16 %0, %2, 1.0, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
          : "=&v"(lo)
          : "v"(x0), "v"(x1), "v"(h));
  return lo;
}

// The mixed precision's tile product on the FP16 matrix cores: every FP32
// operand x is split into a high part h with FP16's precision (B: rounded,
// once per table; A: truncated, cheaper per record) and the FP16 rounding l
// of its remainder (x = h + l to 21-22 bits), and G = Ah Bh + Ah Bl + Al Bh
// (the dropped Al Bl is below 2^-20 of each product) is accumulated in FP32
// by v_mfma_f32_16x16x32_f16, 32 records per instruction.
// B = the xKernel values times 2^(kBExp - eB), with 2^eB above the table's
// largest value; A = v * yKernel, formed in FP32, scaled and split. Each
// image of a 64-record batch is scaled by a power of two that brings its
// largest A below 2^kAExp (power-of-two scalings are exact), and its tile
// product is added to the FP32 sums times the inverse of both scalings.
// Operand layout: lane i + 16 g supplies A[i][8 g + t] and B[8 g + t][i] in
// its eight halves t (the same k order on both sides, so records may take
// any k); D as for gridMfmaKernel. Staging, work queue and output as there.
// GT: the padded table is read from device memory (gkt, gkp: f16TableKernel's
// output) rather than built in LDS, for oversampling factors whose table does
// not fit the LDS (above 221 at K = 7); the values are the same.
template <int NIMG, bool ACCUM, bool ROT, bool GT>
__global__ void __launch_bounds__(kMfmaThreads)
gridMfmaF16Kernel(GridArgs a, const uint32_t* __restrict__ binStart,
                  const uint32_t* __restrict__ extraStart, const uint32_t* __restrict__ itemStart,
                  const uint32_t* __restrict__ itemBin, const uint2* __restrict__ recs,
                  const float2* __restrict__ vis, const float* __restrict__ ktab,
                  const float* __restrict__ gkt, const uint32_t* __restrict__ gkp,
                  float2* __restrict__ layers, float2* __restrict__ partials,
                  uint32_t* __restrict__ queue, int eB) {
#pragma clang fp contract(off)
  constexpr int T = 16, T2 = T * T;
  extern __shared__ float smem[];
  const int K = a.ga.K, OS = a.ga.OS;
  const int rowLen = paddedRow(K, T);
  const int tableWords = GT ? 0 : (OS * rowLen + 3) & ~3;
  const float* kt = GT ? gkt : smem;                                          // FP32
  const uint32_t* kp = GT ? gkp : reinterpret_cast<const uint32_t*>(smem + tableWords);
  int* stPos = reinterpret_cast<int*>(smem + 2 * tableWords);          // [wave][64]
  float2* stVis = reinterpret_cast<float2*>(stPos + kMfmaWaves * 64);  // [wave][64][NIMG]
  if (!GT) {
    const float bScale = ldexpf(1.0f, kBExp - eB);
    uint32_t* ktw = reinterpret_cast<uint32_t*>(smem + tableWords);
    for (int i = threadIdx.x; i < OS * rowLen; i += blockDim.x)
      f16TableEntry(i, ktab, K, rowLen, bScale, smem[i], ktw[i]);
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;  // A row / B column, k group
  int* myPos = stPos + wave * 64;
  float2* myVis = stVis + wave * 64 * NIMG;
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const uint32_t nItems = passItems(binLo, binHi, itemStart);
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  auto take = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(queue, 1u);
    return uint32_t(__builtin_amdgcn_readfirstlane(int(v)));
  };
  uint32_t next = take();
  while (next < nItems) {
    const uint32_t item = next;
    const WorkItem wi = decodeItem(item, binLo, binHi, binStart, itemStart, a.chunk, itemBin);
    next = take();
    const uint32_t bin = wi.bin, start = wi.start, count = wi.count;
    if (ACCUM && count == 0) continue;
    floatx4 acc[2 * NIMG];
#pragma unroll
    for (int e = 0; e < 2 * NIMG; ++e) acc[e] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t b0 = 0; b0 < count; b0 += 64u) {
      // stage 64 records (zero samples past the end: their A rows vanish)
      int pos = 0;
      float2 v[NIMG];
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) v[gi] = make_float2(0.0f, 0.0f);
      const uint32_t it = b0 + uint32_t(lane);
      if (it < count) {
        const uint2 rec = recs[start + it];
        const uint32_t idx = rec.x & 0x3FFFFFFFu;
        const bool conj = (rec.x >> 31) != 0;
        const int kk = int(rec.y >> 12);
        const int ky = kk / OS, kx = kk - ky * OS;
        const int xs = int(rec.y & 63u), ys = int((rec.y >> 6) & 63u);
        pos = (kx * rowLen + (T - 1) + (K - 1) - xs) |
              ((ky * rowLen + (T - 1) + (K - 1) - ys) << 16);
        float2 ph = make_float2(1.0f, 0.0f);
        if (ROT) ph = a.rot[idx];
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          float2 x = weighted(a, vis[size_t(idx) * a.nImgTotal + a.imgs.idx[gi]], idx,
                              a.imgs.idx[gi]);
          if (ROT) x = rotateBy(x, ph);
          if (conj) x.y = -x.y;
          v[gi] = x;
        }
      }
      // each image's scale: 2^e above its largest visibility component in
      // the batch, so that |v yKernel| 2^(kAExp - e - eB) < 2^kAExp (a scale
      // per image keeps images that are power-of-two multiples of each
      // other exact multiples, FP16 subnormals included)
      float up[NIMG];
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) {
        float m = fmaxf(fabsf(v[gi].x), fabsf(v[gi].y));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        int e = 0;
        (void)frexpf(m, &e);  // m < 2^e (0 for m = 0)
        const float down = ldexpf(1.0f, kAExp - e - eB);
        up[gi] = ldexpf(1.0f, e + 2 * eB - kAExp - kBExp);
        v[gi].x = v[gi].x * down;
        v[gi].y = v[gi].y * down;
      }
      myPos[lane] = pos;
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) myVis[lane * NIMG + gi] = v[gi];
      waveSync();
      floatx4 d[2 * NIMG];
#pragma unroll
      for (int c = 0; c < 2 * NIMG; ++c) d[c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int G = 0; G < 2; ++G) {
        int ps[8];
        float kyv[8];
        uint32_t pr[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) ps[t] = myPos[32 * G + 8 * lg + t];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          // B[k][x = li] = xKernel, A[y = li][k] = v * yKernel
          // (yKernelValue * xKernel[i], gridding/wstackinggridder.cpp:574)
          // (unsigned: a y offset of 2^15 or more, OS above 885 at K = 7,
          // sets the word's sign bit)
          pr[t] = kp[(uint32_t(ps[t]) & 0xFFFFu) + li];
          kyv[t] = kt[(uint32_t(ps[t]) >> 16) + li];
        }
        half8 bh, bl;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          bh[t] = __builtin_bit_cast(_Float16, uint16_t(pr[t] & 0xFFFFu));
          bl[t] = __builtin_bit_cast(_Float16, uint16_t(pr[t] >> 16));
        }
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          // per pair of records: the high parts, x rounded toward zero to
          // FP16 (its significand cut to 11 bits), packed by one
          // v_cvt_pkrtz; the low parts, x - h rounded to FP16 (up to
          // 2^-21 |x|), by one v_fma_mix each, written into the two halves
          u32x4 hw[2], lw[2];
#pragma unroll
          for (int t = 0; t < 8; t += 2) {
            const float2 v0 = myVis[(32 * G + 8 * lg + t) * NIMG + gi];
            const float2 v1 = myVis[(32 * G + 8 * lg + t + 1) * NIMG + gi];
            const float xr0 = v0.x * kyv[t], xr1 = v1.x * kyv[t + 1];
            const float xi0 = v0.y * kyv[t], xi1 = v1.y * kyv[t + 1];
            const uint32_t hR = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(xr0, xr1));
            const uint32_t hI = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(xi0, xi1));
            hw[0][t / 2] = hR;
            hw[1][t / 2] = hI;
            lw[0][t / 2] = lowParts(xr0, xr1, hR);
            lw[1][t / 2] = lowParts(xi0, xi1, hI);
          }
          const half8 ah[2] = {__builtin_bit_cast(half8, hw[0]), __builtin_bit_cast(half8, hw[1])};
          const half8 al[2] = {__builtin_bit_cast(half8, lw[0]), __builtin_bit_cast(half8, lw[1])};
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            // the small cross products first, then the high one
            floatx4& dd = d[2 * gi + c];
            dd = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[c], bl, dd, 0, 0, 0);
            dd = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[c], bh, dd, 0, 0, 0);
            dd = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[c], bh, dd, 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 2 * NIMG; ++c) {
        const float u = up[c / 2];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[c][r] = acc[c][r] + d[c][r] * u;
      }
      waveSync();  // the staging is rewritten next batch
    }
    // lane owns column li, rows 4 lg .. 4 lg + 3
    if (wi.j > 0) {
      float2* dst = partials + size_t(extraStart[bin] + wi.j - 1u) * size_t(NIMG * T2);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi)
          dst[gi * T2 + (4 * lg + r) * T + li] = make_float2(acc[2 * gi][r], acc[2 * gi + 1][r]);
    } else {
      const int layer = int(bin / uint32_t(a.tilesPerLayer));
      const int rem = int(bin - uint32_t(layer) * uint32_t(a.tilesPerLayer));
      const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
      const int x = tx * T + li, y0 = ty * T + 4 * lg;
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) {
        float2* base = layers + (size_t(a.imgs.idx[gi]) * a.nLayersPass +
                                 size_t(layer - a.passStart)) * HW;
        const floatx4 re = acc[2 * gi], im = acc[2 * gi + 1];
        if (a.ga.blocked) {
          float4* dst = reinterpret_cast<float4*>(base + cellIndex(x, y0, a.ga.W, 1));
          float4 lo = make_float4(re[0], im[0], re[1], im[1]);
          float4 hi = make_float4(re[2], im[2], re[3], im[3]);
          if (ACCUM) {
            const float4 o0 = dst[0], o1 = dst[1];
            lo.x += o0.x; lo.y += o0.y; lo.z += o0.z; lo.w += o0.w;
            hi.x += o1.x; hi.y += o1.y; hi.z += o1.z; hi.w += o1.w;
          }
          dst[0] = lo;
          dst[1] = hi;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float2* dst = base + cellIndex(x, y0 + r, a.ga.W, 0);
            float2 val = make_float2(re[r], im[r]);
            if (ACCUM) {
              const float2 o = *dst;
              val.x += o.x;
              val.y += o.y;
            }
            *dst = val;
          }
        }
      }
    }
  }
}

// The visibilities of one primary record from the subgrid around its tile
// (sub: the NIMG images of a cell interleaved, row pitch sg cells). KK: the
// kernel size when it is fixed at compile time (0: a.ga.K). The taps are
// summed rows outer, columns inner, as the runtime-K form sums them.
template <int NIMG, int KK>
__device__ __forceinline__ void degridRecord(const GridArgs& a, uint2 rec,
                                             const float* __restrict__ kt,
                                             const float2* __restrict__ sub, int sg,
                                             const double* __restrict__ uvw,
                                             const double* __restrict__ wl,
                                             float2* __restrict__ out) {
  const int K = KK ? KK : a.ga.K, OS = a.ga.OS;
  const uint32_t idx = rec.x & 0x3FFFFFFFu;
  const bool conj = (rec.x >> 31) != 0;
  const int xo = int(rec.y & 63u) - (K - 1), yo = int((rec.y >> 6) & 63u) - (K - 1);
  const int kk = int(rec.y >> 12);
  const int ky = kk / OS, kx = kk - ky * OS;
  float2 acc[NIMG];
#pragma unroll
  for (int gi = 0; gi < NIMG; ++gi) acc[gi] = make_float2(0.0f, 0.0f);
  if constexpr (KK > 0) {
    float xk[KK];
#pragma unroll
    for (int i = 0; i < KK; ++i) xk[i] = kt[kx * KK + i];
    // one row of taps in flight at a time: unrolling the rows hoists every
    // LDS read of the footprint and leaves one wave per SIMD
#pragma nounroll
    for (int j = 0; j < KK; ++j) {
      const float yk = kt[ky * KK + j];
      const float2* row = sub + size_t((yo + j) * sg + xo) * NIMG;
#pragma unroll
      for (int i = 0; i < KK; ++i) {
        // xKernel[i] * yKernelValue (gridding/wstackinggridder.cpp:668)
        const float kv = xk[i] * yk;
        if constexpr (NIMG % 2 == 0) {
#pragma unroll
          for (int gp = 0; gp < NIMG / 2; ++gp) {
            const float4 gv = reinterpret_cast<const float4*>(row + i * NIMG)[gp];
            acc[2 * gp].x += gv.x * kv;
            acc[2 * gp].y += gv.y * kv;
            acc[2 * gp + 1].x += gv.z * kv;
            acc[2 * gp + 1].y += gv.w * kv;
          }
        } else {
#pragma unroll
          for (int gi = 0; gi < NIMG; ++gi) {
            const float2 gv = row[i * NIMG + gi];
            acc[gi].x += gv.x * kv;
            acc[gi].y += gv.y * kv;
          }
        }
      }
    }
  } else {
    for (int j = 0; j < K; ++j) {
      const float yk = kt[ky * K + j];
      const float2* row = sub + size_t((yo + j) * sg + xo) * NIMG;
      for (int i = 0; i < K; ++i) {
        const float kv = kt[kx * K + i] * yk;
#pragma unroll
        for (int gi = 0; gi < NIMG; ++gi) {
          const float2 gv = row[i * NIMG + gi];
          acc[gi].x += gv.x * kv;
          acc[gi].y += gv.y * kv;
        }
      }
    }
  }
  // the degridding phasor is the conjugate of the gridding one (bit for
  // bit: the argument and both sincos forms are odd in it)
  float2 ph = make_float2(1.0f, 0.0f);
  if (a.ga.rotate) {
    ph = a.rot[idx];
    ph.y = -ph.y;
  }
#pragma unroll
  for (int gi = 0; gi < NIMG; ++gi) {
    // folded samples return the sample, others its conjugate (:681-684)
    float2 v = conj ? acc[gi] : make_float2(acc[gi].x, -acc[gi].y);
    if (a.ga.rotate) v = rotateBy(v, ph);
    out[size_t(idx) * a.nImgTotal + a.imgs.idx[gi]] = v;
  }
}

constexpr int kDegridQueue = 128;  // staged primary records per wave
// The (layer, tile) item's subgrid position.
struct DegridItem {
  WorkItem wi;
  int layer, x0, y0;
};

__device__ __forceinline__ DegridItem degridItem(const GridArgs& a, uint32_t item,
                                                 uint32_t binLo, uint32_t binHi,
                                                 const uint32_t* __restrict__ binStart,
                                                 const uint32_t* __restrict__ itemStart,
                                                 const uint32_t* __restrict__ itemBin) {
  DegridItem d;
  d.wi = decodeItem(item, binLo, binHi, binStart, itemStart, a.chunk, itemBin);
  d.layer = int(d.wi.bin / uint32_t(a.tilesPerLayer));
  const int rem = int(d.wi.bin - uint32_t(d.layer) * uint32_t(a.tilesPerLayer));
  const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
  d.x0 = tx * a.ga.tileSize;
  d.y0 = ty * a.ga.tileSize;
  return d;
}

// Layer value of subgrid cell `cell` of item d, image slot gi.
__device__ __forceinline__ float2 subgridCell(const GridArgs& a, const DegridItem& d, int cell,
                                              int gi, const float2* __restrict__ layers) {
  const int sg = a.sg;
  const int cy = cell / sg, cx = cell - cy * sg;
  const int gx = (d.x0 + cx) % a.ga.W, gy = (d.y0 + cy) % a.ga.H;
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  return layers[(size_t(a.imgs.idx[gi]) * a.nLayersPass + (d.layer - a.passStart)) * HW +
                cellIndex(gx, gy, a.ga.W, a.ga.blocked)];
}

// The primary records of one item, degridded from its subgrid `sub` in LDS:
// wave w reads records w * 64 + lane + k * blockDim; qn (wave-uniform)
// primary records wait in the wave's queue, at most 63 between batches.
template <int NIMG, int KK>
__device__ __forceinline__ void degridItemRecords(const GridArgs& a, const WorkItem& wi,
                                                  const uint2* __restrict__ recs,
                                                  const float* kt, const float2* sub,
                                                  uint2* queue, const double* __restrict__ uvw,
                                                  const double* __restrict__ wl,
                                                  float2* __restrict__ out) {
  const int lane = int(threadIdx.x & 63u);
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t start = wi.start, count = wi.count;
  int qn = 0;
  for (uint32_t b = threadIdx.x & ~63u; b < count; b += blockDim.x) {
    const uint32_t it = b + uint32_t(lane);
    uint2 rec = make_uint2(0u, 0u);
    bool primary = false;
    if (it < count) {
      rec = recs[start + it];
      primary = ((rec.x >> 30) & 1u) != 0u;
    }
    const unsigned long long mask = __ballot(primary);
    if (primary) queue[qn + __popcll(mask & below)] = rec;
    qn += __popcll(mask);
    if (qn >= 64) {
      waveSync();
      const uint2 mine = queue[lane];
      const int left = qn - 64;
      uint2 tail = make_uint2(0u, 0u);
      if (lane < left) tail = queue[64 + lane];
      waveSync();
      if (lane < left) queue[lane] = tail;
      qn = left;
      degridRecord<NIMG, KK>(a, mine, kt, sub, a.sg, uvw, wl, out);
    }
  }
  waveSync();
  if (lane < qn) degridRecord<NIMG, KK>(a, queue[lane], kt, sub, a.sg, uvw, wl, out);
}

// Degridding of one image kind, one work item per workgroup at a time: the
// (T + K - 1)^2 subgrid around the item's (layer, tile) is gathered from the
// layer into LDS. Only the primary record of a footprint (the record of the
// tile holding the footprint start) is degridded, so each wave first
// compacts the primary records of its share of the item into a queue in
// LDS and degrids them 64 at a time, one record per lane: no lane idles on
// the secondary records of split footprints.
template <int NIMG, int KK>
__global__ void __launch_bounds__(kThreads)
degridTilesKernel(GridArgs a, const uint32_t* __restrict__ binStart,
                  const uint32_t* __restrict__ itemStart, const uint32_t* __restrict__ itemBin,
                  const uint2* __restrict__ recs, const float* __restrict__ ktab,
                  const __half2* __restrict__ khalf, const double* __restrict__ uvw,
                  const double* __restrict__ wl, const float2* __restrict__ layers,
                  float2* __restrict__ out) {
  extern __shared__ float smem[];
  const int K = KK ? KK : a.ga.K, OS = a.ga.OS, sg2 = a.sg * a.sg;
  float* kt = smem;
  float2* sub = reinterpret_cast<float2*>(smem + ((OS * K + 3) & ~3));  // [sg2][NIMG]
  uint2* queue = reinterpret_cast<uint2*>(sub + size_t(NIMG) * sg2) +
                 (threadIdx.x >> 6) * kDegridQueue;
  loadKernelTable(kt, ktab, khalf, OS * K);
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const uint32_t nItems = passItems(binLo, binHi, itemStart);
  for (uint32_t item = blockIdx.x; item < nItems; item += gridDim.x) {
    // every lane decodes the item (uniform loads)
    const DegridItem d = degridItem(a, item, binLo, binHi, binStart, itemStart, itemBin);
    if (d.wi.count == 0) continue;  // block-uniform
    __syncthreads();  // previous item's readers are done with sub and the queues
    for (int cell = threadIdx.x; cell < sg2; cell += blockDim.x)
#pragma unroll
      for (int gi = 0; gi < NIMG; ++gi) sub[cell * NIMG + gi] = subgridCell(a, d, cell, gi, layers);
    __syncthreads();
    degridItemRecords<NIMG, KK>(a, d.wi, recs, kt, sub, queue, uvw, wl, out);
  }
}


constexpr int kReduceParts = 4;  // threads summing one cell's partial tiles

// Adds the partial tiles of the extra items of every split bin of a pass to
// the layer, in a fixed summation order. Task (e, group): if e is the first
// extra item of its bin, the workgroup sums that bin's partial tiles for one
// group of 256 / kReduceParts cells, kReduceParts threads per cell
// (coalesced over the cells of each partial tile): the dense uv centre's
// bins hold thousands of extra items, and one thread per cell left their
// sums latency-bound.
template <int NIMG>
__global__ void __launch_bounds__(kThreads)
reducePartialsKernel(GridArgs a, const uint32_t* __restrict__ extraStart,
                     const float2* __restrict__ partials, float2* __restrict__ layers) {
  __shared__ float2 red[kReduceParts][kThreads / kReduceParts];
  const uint32_t binLo = uint32_t(a.layerLo) * uint32_t(a.tilesPerLayer);
  const uint32_t binHi = uint32_t(a.layerHi) * uint32_t(a.tilesPerLayer);
  const uint32_t eLo = extraStart[binLo], eHi = extraStart[binHi];
  const int T = a.ga.tileSize, T2 = T * T;
  constexpr int kCells = kThreads / kReduceParts;  // cells per task
  const uint32_t groups = uint32_t((T2 + kCells - 1) / kCells);
  const size_t HW = size_t(a.ga.W) * a.ga.H;
  const size_t tasks = size_t(eHi - eLo) * groups;
  const int part = int(threadIdx.x) / kCells, cl = int(threadIdx.x) - part * kCells;
  for (size_t task = blockIdx.x; task < tasks; task += gridDim.x) {
    const uint32_t e = eLo + uint32_t(task / groups);
    const uint32_t grp = uint32_t(task - size_t(e - eLo) * groups);
    const uint32_t bin = ownerOf(extraStart, binLo, binHi, e);
    const uint32_t first = extraStart[bin];
    if (e != first) continue;  // another task of the bin does the work (workgroup-uniform)
    const int c = int(grp) * kCells + cl;
    const bool valid = c < T2;
    const uint32_t nEx = extraStart[bin + 1] - first;
    const int layer = int(bin / uint32_t(a.tilesPerLayer));
    const int rem = int(bin - uint32_t(layer) * uint32_t(a.tilesPerLayer));
    const int ty = rem / a.ga.tilesX, tx = rem - ty * a.ga.tilesX;
    const int cy = c / T, cx = c - cy * T;
#pragma unroll
    for (int gi = 0; gi < NIMG; ++gi) {
      // part p of a cell sums the partial tiles p, p + P, p + 2P, ... (P
      // parts) with four independent sums; the parts are then added in
      // part order
      float2 s4[4] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
      if (valid) {
        const float2* src = partials + (size_t(first) * NIMG + gi) * T2 + c;
        const size_t stride = size_t(NIMG) * T2;
        uint32_t q = uint32_t(part);
        for (; q + 3u * kReduceParts < nEx; q += 4u * kReduceParts) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const float2 v = src[size_t(q + uint32_t(k * kReduceParts)) * stride];
            s4[k].x += v.x;
            s4[k].y += v.y;
          }
        }
        for (; q < nEx; q += kReduceParts) {
          const float2 v = src[size_t(q) * stride];
          s4[0].x += v.x;
          s4[0].y += v.y;
        }
      }
      red[part][cl] = make_float2((s4[0].x + s4[1].x) + (s4[2].x + s4[3].x),
                                  (s4[0].y + s4[1].y) + (s4[2].y + s4[3].y));
      __syncthreads();
      if (part == 0 && valid) {
        float2 t = red[0][cl];
#pragma unroll
        for (int pp = 1; pp < kReduceParts; ++pp) {
          t.x += red[pp][cl].x;
          t.y += red[pp][cl].y;
        }
        const size_t off = cellIndex(tx * T + cx, ty * T + cy, a.ga.W, a.ga.blocked);
        float2* dst = layers +
                      (size_t(a.imgs.idx[gi]) * a.nLayersPass + (layer - a.passStart)) * HW + off;
        const float2 o = *dst;
        *dst = make_float2(o.x + t.x, o.y + t.y);
      }
      __syncthreads();  // red is rewritten by the next image or task
    }
  }
}

// ----------------------------------------------------- plain per-channel
struct PlainArgs {
  int W, H, K, OS, mid;
  float dl, dm;
  int nW;
  float minW, maxW;
  int isComplex;
  int layerLo, layerHi, passStart, nLayersPass;
  int img, nImgTotal;
  int rotate;
  float lShift, mShift;
  int ch;
  size_t nrow, nchan;
  int blocked;
  const float* cw;  // as GridArgs
};

struct PlainPlacement {
  int ok;
  int layer, x, y, kx, ky, conj;
  float rotRe, rotIm;
};

__device__ __forceinline__ PlainPlacement placePlain(const PlainArgs& a,
                                                     const double* __restrict__ uvw,
                                                     const double* __restrict__ wl,
                                                     size_t r, float rotSign) {
  PlainPlacement p;
  p.ok = 0;
  const float lambda = float(wl[a.ch]);
  const float um = float(uvw[r * 3]), vm = float(uvw[r * 3 + 1]),
              wm = float(uvw[r * 3 + 2]);
  float u = um / lambda, v = vm / lambda, w = wm / lambda;
  p.rotRe = 1.0f;
  p.rotIm = 0.0f;
  if (a.rotate) {
    const float rad = rotSign * 2.0f * float(kPi) * (um * a.lShift + vm * a.mShift) / lambda;
    sincosf(rad, &p.rotIm, &p.rotRe);
  }
  p.conj = 0;
  if (w < 0.0f && !a.isComplex) {
    u = -u;
    v = -v;
    w = -w;
    p.conj = 1;
  }
  int layer = 0;
  if (a.nW != 1) {
    float lf = a.isComplex ? roundf((w + a.maxW) * (a.nW - 1) / (a.maxW + a.maxW))
                           : roundf((fabsf(w) - a.minW) * (a.nW - 1) / (a.maxW - a.minW));
    if (!(lf >= 0.0f) || lf >= float(a.nW)) return p;
    layer = int(lf);
  }
  if (layer < a.layerLo || layer >= a.layerHi) return p;
  const float xE = u * a.dl * float(a.W), yE = v * a.dm * float(a.H);
  if (!(fabsf(xE) < 1.0e9f) || !(fabsf(yE) < 1.0e9f)) return p;
  int x = int(roundf(xE)), y = int(roundf(yE));
  int kx = int(roundf((xE - float(x)) * a.OS));
  int ky = int(roundf((yE - float(y)) * a.OS));
  kx = (kx + (a.OS * 3) / 2) % a.OS;
  ky = (ky + (a.OS * 3) / 2) % a.OS;
  if (!(x > -a.W / 2 && y > -a.H / 2 && x <= a.W / 2 && y <= a.H / 2)) return p;
  if (x < 0) x += a.W;
  if (y < 0) y += a.H;
  p.ok = 1;
  p.layer = layer;
  p.x = x;
  p.y = y;
  p.kx = kx;
  p.ky = ky;
  return p;
}

__global__ void __launch_bounds__(kThreads)
gridPlainKernel(PlainArgs a, const double* __restrict__ uvw,
                const double* __restrict__ wl, const float2* __restrict__ vis,
                const float* __restrict__ ktab, float2* __restrict__ layers) {
  const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= a.nrow) return;
  const PlainPlacement p = placePlain(a, uvw, wl, r, -1.0f);
  if (!p.ok) return;
  float2 v = vis[(r * a.nchan + a.ch) * a.nImgTotal + a.img];
  if (a.cw) {
    const float w = a.cw[(r * a.nchan + a.ch) * a.nImgTotal + a.img];
    v = make_float2(v.x * w, v.y * w);
  }
  if (a.rotate) v = make_float2(v.x * p.rotRe - v.y * p.rotIm, v.x * p.rotIm + v.y * p.rotRe);
  if (p.conj) v.y = -v.y;
  const size_t HW = size_t(a.W) * a.H;
  float2* layer = layers + (size_t(a.img) * a.nLayersPass + (p.layer - a.passStart)) * HW;
  for (int j = 0; j < a.K; ++j) {
    const float yk = ktab[p.ky * a.K + j];
    int cy = p.y + j - a.mid;
    if (cy < 0) cy += a.H;
    if (cy >= a.H) cy -= a.H;
    for (int i = 0; i < a.K; ++i) {
      int cx = p.x + i - a.mid;
      if (cx < 0) cx += a.W;
      if (cx >= a.W) cx -= a.W;
      const float kv = yk * ktab[p.kx * a.K + i];
      float2* dst = layer + cellIndex(cx, cy, a.W, a.blocked);
      unsafeAtomicAdd(&dst->x, v.x * kv);
      unsafeAtomicAdd(&dst->y, v.y * kv);
    }
  }
}

__global__ void __launch_bounds__(kThreads)
degridPlainKernel(PlainArgs a, const double* __restrict__ uvw,
                  const double* __restrict__ wl, const float* __restrict__ ktab,
                  const float2* __restrict__ layers, float2* __restrict__ out) {
  const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= a.nrow) return;
  const PlainPlacement p = placePlain(a, uvw, wl, r, 1.0f);
  float2* dst = out + (r * a.nchan + a.ch) * a.nImgTotal + a.img;
  if (!p.ok) return;  // stays NaN (filled before the first pass)
  const size_t HW = size_t(a.W) * a.H;
  const float2* layer = layers + (size_t(a.img) * a.nLayersPass + (p.layer - a.passStart)) * HW;
  float2 acc = make_float2(0.0f, 0.0f);
  for (int j = 0; j < a.K; ++j) {
    const float yk = ktab[p.ky * a.K + j];
    int cy = p.y + j - a.mid;
    if (cy < 0) cy += a.H;
    if (cy >= a.H) cy -= a.H;
    for (int i = 0; i < a.K; ++i) {
      int cx = p.x + i - a.mid;
      if (cx < 0) cx += a.W;
      if (cx >= a.W) cx -= a.W;
      const float kv = ktab[p.kx * a.K + i] * yk;
      const float2 gv = layer[cellIndex(cx, cy, a.W, a.blocked)];
      acc.x += gv.x * kv;
      acc.y += gv.y * kv;
    }
  }
  float2 v = p.conj ? acc : make_float2(acc.x, -acc.y);
  if (a.rotate) v = make_float2(v.x * p.rotRe - v.y * p.rotIm, v.x * p.rotIm + v.y * p.rotRe);
  *dst = v;
}

// ------------------------------------------------------ image domain
struct ImgArgs {
  int W, H;
  double dl, dm, lShift, mShift;
  int precision;
  LayerMap maps[2];  // real, complex
  int isComplex[kMaxImages];
  int passStart, nLayersInPass, nLayersPass;
  int blocked;
  int fresh;  // the accumulators start from zero (not read)
};


// Invert: every layer of the pass onto the image, FFT-ordered pixel p.
// projectOnImageAndCorrect (gridding/wstackinggridder.cpp:874-907) for all
// layers of the pass at once.
__global__ void __launch_bounds__(kThreads)
projectKernel(ImgArgs a, const float2* __restrict__ layers,
              float* __restrict__ sum, float* __restrict__ comp,
              float* __restrict__ isum, float* __restrict__ icomp) {
  const size_t HW = size_t(a.W) * a.H;
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  const int img = blockIdx.y;
  const int cplx = a.isComplex[img];
  int x, y;
  cellOf(p, a.W, a.blocked, x, y);
  int ySrc = (a.H - y) + a.H / 2;
  if (ySrc >= a.H) ySrc -= a.H;
  int xSrc = x + a.W / 2;
  if (xSrc >= a.W) xSrc -= a.W;
  const double m = pixelM(ySrc, a.H, a.dm, a.mShift);
  const double l = pixelL(xSrc, a.W, a.dl, a.lShift);
  double np = nPrime(l, m);
  if (a.precision == HWS_PRECISION_REFERENCE) np = double(float(np));
  const size_t o = size_t(img) * HW + size_t(ySrc) * a.W + xSrc;
  float s = 0.0f, c = 0.0f, is = 0.0f, ic = 0.0f;
  if (!a.fresh) {
    s = sum[o];
    c = comp[o];
    if (cplx) {
      is = isum[o];
      ic = icomp[o];
    }
  }
  const LayerMap& map = a.maps[cplx];
  const float2* src = layers + size_t(img) * a.nLayersPass * HW + p;
  for (int li = 0; li < a.nLayersInPass; ++li) {
    // (streamed once: a non-temporal load)
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 fv = __builtin_nontemporal_load(reinterpret_cast<const f2*>(src + size_t(li) * HW));
    const float2 F = make_float2(fv.x, fv.y);
    const double w = layerToW(map, a.passStart + li);
    const float2 t = a.precision == HWS_PRECISION_REFERENCE ? projectTerm<true>(F, w, np)
                                                            : projectTerm<false>(F, w, np);
    kahanAdd(s, c, t.x);
    if (cplx) kahanAdd(is, ic, t.y);
  }
  sum[o] = s;
  comp[o] = c;
  // (a fresh pass also zeroes the imaginary accumulators of real images)
  if (cplx || (a.fresh && isum)) {
    isum[o] = is;
    icomp[o] = ic;
  }
}

// Plain form: one layer, FP32 phase from the float n' table.
__global__ void __launch_bounds__(kThreads)
projectPlainKernel(int W, int H, int blocked, const float2* __restrict__ layer,
                   const float* __restrict__ ntab, float w, int cplx,
                   float* __restrict__ sum, float* __restrict__ isum) {
  const size_t HW = size_t(W) * H;
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  int x, y;
  cellOf(p, W, blocked, x, y);
  int ySrc = (H - y) + H / 2;
  if (ySrc >= H) ySrc -= H;
  int xSrc = x + W / 2;
  if (xSrc >= W) xSrc -= W;
  const size_t o = size_t(ySrc) * W + xSrc;
  const float2 F = layer[p];
  float s, c;
  sincosf(-2.0f * float(kPi) * w * ntab[p], &s, &c);
  sum[o] += F.x * c - F.y * s;
  if (cplx) isum[o] += F.x * s + F.y * c;
}

// n' tables as the reference stores them (T = float): order 0 for invert
// (cpp:851-872), 1 for prediction (cpp:909-932).
__global__ void nTableKernel(int W, int H, int blocked, double dl, double dm,
                             double ls, double ms, int order, float* __restrict__ out) {
  const size_t HW = size_t(W) * H;
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  int x, y;
  cellOf(p, W, blocked, x, y);
  int xs, ys;
  if (order == 0) {
    ys = (H - y) + H / 2;
    if (ys >= H) ys -= H;
    xs = x + W / 2;
    if (xs >= W) xs -= W;
  } else {
    ys = y + H / 2;
    if (ys >= H) ys -= H;
    xs = (W - x) + W / 2;
    if (xs >= W) xs -= W;
  }
  const double m = pixelM(ys, H, dm, ms);
  const double l = pixelL(xs, W, dl, ls);
  out[p] = float(nPrime(l, m));
}

// FinalizeImage (cpp:698-715) + correctImageForKernel<false> for T=float
// (cpp:773-827): image = ((sum - comp) * mult) / (Dx * Dy / OS^2).
__global__ void __launch_bounds__(kThreads)
finalizeKernel(int W, int H, const float* __restrict__ sum,
               const float* __restrict__ comp, double mult,
               const float* __restrict__ corrX, const float* __restrict__ corrY,
               double norm, int correct, const double* __restrict__ weightSums,
               float* __restrict__ out) {
  const size_t HW = size_t(W) * H;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= HW) return;
  const size_t img = blockIdx.y;
  const int x = int(i % W), y = int(i / W);
  float v = sum[img * HW + i] - comp[img * HW + i];
  // weighted contexts normalise each image by its own weight sum, as
  // WSMSGridder::Invert finalizes with 1 / totalWeight (wsmsgridder.cpp:436)
  const double m = weightSums ? mult / weightSums[img * 4] : mult;
  v = float(double(v) * m);
  if (correct) {
    const float xv = corrX[x >= W / 2 ? x - W / 2 : W / 2 - x];
    const float yv = corrY[y >= H / 2 ? y - H / 2 : H / 2 - y];
    v = float(double(v) / (double(xv * yv) * norm));
  }
  out[img * HW + i] = v;
}

// initializePrediction (cpp:829-849): horizon mask then divide by the
// kernel correction.
__global__ void __launch_bounds__(kThreads)
initPredictionKernel(int W, int H, double dl, double dm, double ls, double ms,
                     const float* __restrict__ in, float* __restrict__ model,
                     const float* __restrict__ corrX,
                     const float* __restrict__ corrY, double norm, int correct) {
  const size_t HW = size_t(W) * H;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= HW) return;
  const size_t img = blockIdx.y;
  const int x = int(i % W), y = int(i / W);
  const double m = pixelM(y, H, dm, ms);
  const double l = pixelL(x, W, dl, ls);
  float v = model[img * HW + i];
  v = (isfinite(v) && nPrimeInside(l, m)) ? in[img * HW + i] : 0.0f;
  if (correct) {
    const float xv = corrX[x >= W / 2 ? x - W / 2 : W / 2 - x];
    const float yv = corrY[y >= H / 2 ? y - H / 2 : H / 2 - y];
    v = float(double(v) / (double(xv * yv) * norm));
  }
  model[img * HW + i] = v;
}

// copyImageToLayerAndInverseCorrect (cpp:934-969) for every layer of the
// pass; layer-ordered pixel p reads model pixel (xDest, yDest).
__global__ void __launch_bounds__(kThreads)
imageToLayersKernel(ImgArgs a, const float* __restrict__ modelRe,
                    const float* __restrict__ modelIm,
                    float2* __restrict__ layers) {
  const size_t HW = size_t(a.W) * a.H;
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  const int img = blockIdx.y;
  const int cplx = a.isComplex[img];
  int x, y;
  cellOf(p, a.W, a.blocked, x, y);
  int yDest = y + a.H / 2;
  if (yDest >= a.H) yDest -= a.H;
  int xDest = (a.W - x) + a.W / 2;
  if (xDest >= a.W) xDest -= a.W;
  const double m = pixelM(yDest, a.H, a.dm, a.mShift);
  const double l = pixelL(xDest, a.W, a.dl, a.lShift);
  double np = nPrimeSeries(l, m);  // as the forward transform's fused row pass
  if (a.precision == HWS_PRECISION_REFERENCE) np = double(float(np));
  const size_t src = size_t(img) * HW + size_t(yDest) * a.W + xDest;
  const float re = modelRe[src];
  const float imv = cplx ? -modelIm[src] : 0.0f;
  const LayerMap& map = a.maps[cplx];
  float2* dst = layers + size_t(img) * a.nLayersPass * HW + p;
  for (int li = 0; li < a.nLayersInPass; ++li) {
    const double w = layerToW(map, a.passStart + li);
    dst[size_t(li) * HW] = a.precision == HWS_PRECISION_REFERENCE
                               ? modelTerm<true>(re, imv, cplx != 0, w, np)
                               : modelTerm<false>(re, imv, cplx != 0, w, np);
  }
}

// Plain form: one layer of one image from the float predict-order n' table.
__global__ void __launch_bounds__(kThreads)
imageToLayerPlainKernel(int W, int H, int blocked, const float* __restrict__ modelRe,
                        const float* __restrict__ modelIm,
                        const float* __restrict__ ntab, float w, int cplx,
                        float2* __restrict__ layer) {
  const size_t HW = size_t(W) * H;
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= HW) return;
  int x, y;
  cellOf(p, W, blocked, x, y);
  int yDest = y + H / 2;
  if (yDest >= H) yDest -= H;
  int xDest = (W - x) + W / 2;
  if (xDest >= W) xDest -= W;
  const size_t src = size_t(yDest) * W + xDest;
  float s, c;
  sincosf(2.0f * float(kPi) * w * ntab[p], &s, &c);
  const float re = modelRe[src];
  if (cplx) {
    const float imv = -modelIm[src];
    layer[p] = make_float2(re * c + imv * s, imv * c - re * s);
  } else {
    layer[p] = make_float2(re * c, -re * s);
  }
}

template <typename T>
hws_status ensureAlloc(hws_gridder* g, T*& ptr, size_t& cap, size_t n,
                       const char* what) {
  if (cap >= n && ptr) return HWS_OK;
  if (ptr) {
    (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
  const hws_status st = checkHip(g, hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T)), what);
  if (st == HWS_OK) cap = n;
  return st;
}

unsigned blocksFor(size_t n, int threads = kThreads) {
  return unsigned((n + threads - 1) / threads);
}

}  // namespace

// Exclusive scan of n values into out[0..n] (out[n] = total); scratch holds
// the block sums of every level (scanScratchSize(n) values).
size_t scanScratchSize(size_t n) {
  size_t total = 0;
  while (n > size_t(kScanBlock)) {
    const size_t nb = (n + kScanBlock - 1) / kScanBlock;
    total += 2 * (nb + 1);
    n = nb;
  }
  return total + 2 * 2 + 64;
}

hws_status exclusiveScan(hws_gridder* g, const uint32_t* in, uint32_t* out,
                         size_t n, uint32_t* scratch) {
  if (n == 0) {
    hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, g->stream, out, 0u, size_t(1));
    return checkHip(g, hipGetLastError(), "scan fill");
  }
  const size_t nblocks = (n + kScanBlock - 1) / kScanBlock;
  uint32_t* sums = scratch;
  uint32_t* offs = scratch + nblocks + 1;
  hipLaunchKernelGGL(scanBlocksKernel, dim3(unsigned(nblocks)), dim3(kThreads), 0, g->stream, in, out, n, sums);
  HWS_TRY(checkHip(g, hipGetLastError(), "scan blocks"));
  if (nblocks > 1) {
    HWS_TRY(exclusiveScan(g, sums, offs, nblocks, scratch + 2 * (nblocks + 1)));
    hipLaunchKernelGGL(scanAddKernel, dim3(blocksFor(n)), dim3(kThreads), 0, g->stream, out, n, offs);
    HWS_TRY(checkHip(g, hipGetLastError(), "scan add"));
  }
  hipLaunchKernelGGL(scanTotalKernel, dim3(1), dim3(64), 0, g->stream, in, out, n);
  return checkHip(g, hipGetLastError(), "scan total");
}

namespace {

GeomArgs geomArgs(const hws_gridder* g) {
  GeomArgs ga;
  ga.W = g->W;
  ga.H = g->H;
  ga.K = g->K;
  ga.OS = g->OS;
  ga.mid = g->K / 2;
  ga.dl = g->cfg.pixel_size_x;
  ga.dm = g->cfg.pixel_size_y;
  ga.tileSize = g->plan.tileSize;
  ga.tilesX = g->plan.tilesX;
  ga.tilesY = g->plan.tilesY;
  ga.tileShift = -1;
  for (int sh = 0; sh < 8; ++sh)
    if ((1 << sh) == ga.tileSize) ga.tileShift = sh;
  ga.rotate = g->cfg.rotate_visibilities &&
              (g->cfg.l_shift != 0.0 || g->cfg.m_shift != 0.0);
  ga.lShift = g->cfg.l_shift;
  ga.mShift = g->cfg.m_shift;
  ga.fastPhase = g->cfg.precision == HWS_PRECISION_MIXED;
  ga.blocked = g->blocked;
  return ga;
}

ImgArgs imgArgs(const hws_gridder* g, int pass) {
  ImgArgs a;
  a.W = g->W;
  a.H = g->H;
  a.dl = g->cfg.pixel_size_x;
  a.dm = g->cfg.pixel_size_y;
  a.lShift = g->cfg.l_shift;
  a.mShift = g->cfg.m_shift;
  a.precision = g->cfg.precision;
  a.maps[0] = layerMap(g, 0);
  a.maps[1] = layerMap(g, 1);
  for (int i = 0; i < kMaxImages; ++i) a.isComplex[i] = (i < g->nImg) ? g->isComplex[i] : 0;
  a.passStart = passLayerStart(g, pass);
  a.nLayersInPass = passLayerStart(g, pass + 1) - a.passStart;
  a.nLayersPass = g->maxLayersPerPass;
  a.blocked = g->blocked;
  a.fresh = 0;
  return a;
}

// The matrix-core grid kernel covers 16 x 16 tiles with the padded table.
bool mfmaKernelUsable(const hws_gridder* g) {
  return g->plan.tileSize == 16 && cellTablePadded(g->K, g->OS, 16) && !g->noMfma;
}

// The padded table of the FP16 matrix-core kernel too large for its LDS, but
// indexable by its 16-bit record offsets: read from device memory.
bool f16TableGlobal(const hws_gridder* g) {
  return !cellTablePadded(g->K, g->OS, 16) && size_t(g->OS) * paddedRow(g->K, 16) < 65536;
}

// The mixed precision grids on the FP16 matrix cores (split operands).
bool mfmaF16(const hws_gridder* g) {
  return g->plan.tileSize == 16 && !g->noMfma && g->cfg.precision == HWS_PRECISION_MIXED &&
         !g->noF16Mfma && (cellTablePadded(g->K, g->OS, 16) || f16TableGlobal(g));
}

size_t gridLdsBytes(const hws_gridder* g, int nimg, bool degrid) {
  const int T = g->tileSize;
  const size_t kt = size_t((g->OS * g->K + 3) & ~3) * sizeof(float);
  if (degrid) {
    const int sg = T + g->K - 1;
    return kt + size_t(nimg) * sg * sg * sizeof(float2) +
           size_t(kThreads / 64) * kDegridQueue * sizeof(uint2);
  }
  if (mfmaKernelUsable(g) || mfmaF16(g)) {
    // the FP32 table, and for the FP16 kernel the (hi, lo) pairs beside it
    // (none where that kernel reads them from device memory)
    size_t table = size_t((g->OS * paddedRow(g->K, T) + 3) & ~3) * sizeof(float) *
                   (mfmaF16(g) ? 2 : 1);
    if (mfmaF16(g) && f16TableGlobal(g)) table = 0;
    return table + kMfmaWaves * 64 * (sizeof(int) + size_t(nimg) * sizeof(float2));
  }
  if (cellKernelTile(T)) {
    // the cell kernel keeps its tile in registers; LDS holds the table only
    if (cellTablePadded(g->K, g->OS, T))
      return size_t(g->OS) * paddedRow(g->K, T) * sizeof(float);
    return kt;
  }
  // per wave: a private T x pitch tile per image component + 64 staged records
  const size_t perWave = size_t(nimg) * 2 * T * privatePitch(T) * sizeof(float) +
                         64 * (sizeof(int2) + size_t(nimg) * sizeof(float2));
  return kt + kGridWaves * perWave;
}

// LDS a workgroup may use (gfx950: 160 KiB per CU).
size_t maxLdsBytes(const hws_gridder* g) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, g->cfg.device) !=
          hipSuccess ||
      v <= 0)
    v = 64 * 1024;
  return size_t(v);
}

int persistentBlocks(const hws_gridder* g, int perCu) {
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->cfg.device);
  return std::max(1, cus) * perCu;
}

// Records per work item: about a sixteenth of an even share of the records
// (about 1.2 per sample) over 4 workgroups per CU, a power of two in
// [1024, 8192] (measured best: 4096-8192 at the 4096^2 benchmark); or the
// configured value.
uint32_t workChunk(const hws_gridder* g, size_t nSamples) {
  if (g->cfg.work_chunk) return uint32_t(g->cfg.work_chunk);
  const size_t target = (nSamples + nSamples / 4) / (size_t(persistentBlocks(g, 4)) * 16) + 1;
  size_t c = 1024;
  while (c < target && c < 8192) c <<= 1;
  return uint32_t(c);
}

}  // namespace

hws_status buildF16Table(hws_gridder* g) {
  if (!f16TableGlobal(g) || g->cfg.precision != HWS_PRECISION_MIXED) return HWS_OK;
  const size_t n = size_t(g->OS) * paddedRow(g->K, 16);
  HWS_TRY(checkHip(g, hipMalloc(&g->dF16Table, 2 * n * sizeof(float)), "padded kernel table"));
  hipLaunchKernelGGL(f16TableKernel, dim3(unsigned((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, g->stream, g->dKernel, g->K, g->OS, g->kernelExp,
                     g->dF16Table, reinterpret_cast<uint32_t*>(g->dF16Table + n));
  HWS_TRY(checkHip(g, hipGetLastError(), "padded kernel table launch"));
  return checkHip(g, hipStreamSynchronize(g->stream), "padded kernel table");
}


hws_status launchFill(hws_gridder* g, void* ptr, uint32_t value, size_t n32) {
  if (n32 == 0) return HWS_OK;
  const unsigned blocks = unsigned(std::min<size_t>(blocksFor(n32), 65536));
  hipLaunchKernelGGL(fillKernel, dim3(blocks), dim3(kThreads), 0, g->stream,
                     static_cast<uint32_t*>(ptr), value, n32);
  return checkHip(g, hipGetLastError(), "fill launch");
}

hws_status buildPlan(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl) {
  Plan& pl = g->plan;
  pl.valid = false;
  pl.device = false;  // hws_plan_device marks its own plans
  ++g->generation->value;
  // imaging weights belong to the planned device path: planWeights sets them
  // up after this plan, the host protocol grids its staged rows unweighted
  pl.weighted = false;
  g->weightSumsValid = false;
  const size_t n = nrow * nchan;
  if (n >= (size_t(1) << 30))
    return setLastError(g, HWS_ERR_UNSUPPORTED,
                        "a plan holds at most 2^30 - 1 (row, channel) samples");
  pl.nrow = nrow;
  pl.nchan = nchan;
  pl.uvw = dUvw;
  pl.wavelengths = dWl;
  pl.tileSize = g->tileSize;
  pl.tilesX = g->W / pl.tileSize;
  pl.tilesY = g->H / pl.tileSize;
  const size_t tpl = size_t(pl.tilesX) * pl.tilesY;
  pl.nbins = tpl * size_t(g->nW);
  if (pl.nbins >= 0xFFFFFFFFull)
    return setLastError(g, HWS_ERR_UNSUPPORTED, "too many (layer, tile) bins");
  HWS_TRY(ensureAlloc(g, pl.scanScratch, pl.scanScratchCapacity, scanScratchSize(pl.nbins),
                      "plan scan scratch"));
  if (!pl.counters)
    HWS_TRY(checkHip(g, hipMalloc(&pl.counters, kPlanCounters * sizeof(unsigned long long)),
                     "plan counters"));
  if (!pl.queue) HWS_TRY(checkHip(g, hipMalloc(&pl.queue, 64 * sizeof(uint32_t)), "plan queue"));
  HWS_TRY(launchFill(g, pl.counters, 0u, 2 * kPlanCounters));

  const GeomArgs ga = geomArgs(g);
  pl.chunk = workChunk(g, n);
  // rows in the order of their middle channel's bin (the layer map of the
  // real images if there are any)
  int keyKind = 1;
  for (int i = 0; i < g->nImg; ++i)
    if (!g->isComplex[i]) keyKind = 0;
  HWS_TRY(ensureAlloc(g, pl.rowKeys, pl.rowKeysCapacity, std::max<size_t>(nrow, 1), "row keys"));
  HWS_TRY(ensureAlloc(g, pl.rowKeysAlt, pl.rowKeysAltCapacity, std::max<size_t>(nrow, 1),
                      "row keys"));
  HWS_TRY(ensureAlloc(g, pl.rowOrder, pl.rowOrderCapacity, std::max<size_t>(nrow, 1),
                      "row order"));
  HWS_TRY(ensureAlloc(g, pl.rowOrderAlt, pl.rowOrderAltCapacity, std::max<size_t>(nrow, 1),
                      "row order"));
  if (n > 0) {
    hipLaunchKernelGGL(rowKeyKernel, dim3(blocksFor(nrow)), dim3(kThreads), 0, g->stream, dUvw,
                       dWl, nrow, int(nchan), ga, layerMap(g, keyKind), pl.rowKeys, pl.rowOrder);
    HWS_TRY(checkHip(g, hipGetLastError(), "row key launch"));
    HWS_TRY(sortRows(g, pl.rowKeys, pl.rowKeysAlt, pl.rowOrder, pl.rowOrderAlt, nrow));
  }
  // row groups of PER * 256 samples, visited with a stride coprime to their
  // number: the count pass takes 16 samples per thread and group, the
  // scatter (which keeps its records in registers until the group's
  // positions are known) 2 (measured best of 2-16 and 1-4 at the 4096^2
  // benchmark)
  constexpr int countPer = 16, scatterPer = 2;
  auto groupsOf = [&](int per, uint32_t& ngroups, uint32_t& step, unsigned& blocks) {
    const size_t gs = size_t(per) * kThreads;
    ngroups = uint32_t((n + gs - 1) / gs);
    step = 1;
    if (ngroups > 2) {
      step = uint32_t(double(ngroups) * 0.6180339887498949);
      auto gcd = [](uint32_t a, uint32_t b) {
        while (b) {
          const uint32_t t = a % b;
          a = b;
          b = t;
        }
        return a;
      };
      while (gcd(step, ngroups) != 1u) ++step;
    }
    blocks = unsigned(std::max<size_t>(1, std::min<size_t>(ngroups, size_t(persistentBlocks(g, 8)))));
  };
  uint32_t cGroups, cStep, sGroups, sStep;
  unsigned cBlocks, sBlocks;
  groupsOf(countPer, cGroups, cStep, cBlocks);
  groupsOf(scatterPer, sGroups, sStep, sBlocks);
  for (int kind = 0; kind < 2; ++kind) {
    KindPlan& kp = pl.kinds[kind];
    kp.kind = kind;
    kp.nImg = 0;
    for (int i = 0; i < g->nImg; ++i)
      if (int(g->isComplex[i]) == kind) kp.img[kp.nImg++] = i;
    kp.active = kp.nImg > 0;
    if (!kp.active) continue;
    HWS_TRY(ensureAlloc(g, kp.counts, kp.countsCapacity, pl.nbins + 1, "plan counts"));
    HWS_TRY(ensureAlloc(g, kp.binStart, kp.binCapacity, pl.nbins + 1, "plan bin offsets"));
    HWS_TRY(ensureAlloc(g, kp.extraStart, kp.extraCapacity, pl.nbins + 1, "plan item offsets"));
    HWS_TRY(ensureAlloc(g, kp.itemStart, kp.itemCapacity, pl.nbins + 1, "plan item offsets"));
    HWS_TRY(launchFill(g, kp.counts, 0u, pl.nbins + 1));
    KindArgs ka{layerMap(g, kind), kp.counts, nullptr, nullptr, pl.counters + 3 * kind,
                nullptr, 0u, nullptr, 0u, pl.counters + 18 + kind};
    if (n > 0) {
#define HWS_COUNT(P)                                                                      \
  hipLaunchKernelGGL((binKernel<false, false, P>), dim3(cBlocks), dim3(kThreads), 0, g->stream, \
                     dUvw, dWl, pl.rowOrder, nrow, int(nchan), cGroups, cStep, ga, ka)
      HWS_COUNT(countPer);
#undef HWS_COUNT
      HWS_TRY(checkHip(g, hipGetLastError(), "bin count launch"));
    }
    HWS_TRY(exclusiveScan(g, kp.counts, kp.binStart, pl.nbins, pl.scanScratch));
    // work items: the counts array is free until the scatter cursor is set
    hipLaunchKernelGGL(extraItemsKernel, dim3(blocksFor(pl.nbins)), dim3(kThreads), 0,
                       g->stream, kp.binStart, pl.nbins, pl.chunk, kp.counts,
                       pl.counters + 6 + kind);
    HWS_TRY(checkHip(g, hipGetLastError(), "work item launch"));
    HWS_TRY(exclusiveScan(g, kp.counts, kp.extraStart, pl.nbins, pl.scanScratch));
    hipLaunchKernelGGL(binItemsKernel, dim3(blocksFor(pl.nbins)), dim3(kThreads), 0, g->stream,
                       kp.binStart, kp.extraStart, pl.nbins, kp.counts);
    HWS_TRY(checkHip(g, hipGetLastError(), "work item launch"));
    HWS_TRY(exclusiveScan(g, kp.counts, kp.itemStart, pl.nbins, pl.scanScratch));
    // the record and item totals size the sorted array and partials
    uint32_t totals[3] = {0, 0, 0};
    unsigned long long placed[3] = {0ull, 0ull, 0ull};
    HWS_TRY(checkHip(g, hipMemcpyAsync(placed, pl.counters + 3 * kind, sizeof(placed),
                                       hipMemcpyDeviceToHost, g->stream), "sample counts"));
    HWS_TRY(checkHip(g, hipMemcpyAsync(&totals[0], kp.binStart + pl.nbins, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, g->stream), "record total"));
    HWS_TRY(checkHip(g, hipMemcpyAsync(&totals[1], kp.extraStart + pl.nbins, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, g->stream), "item total"));
    HWS_TRY(checkHip(g, hipMemcpyAsync(&totals[2], kp.itemStart + pl.nbins, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, g->stream), "item total"));
    HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "plan totals"));
    const uint32_t total = totals[0];
    kp.nSkipped = placed[1] + placed[2];
    HWS_TRY(ensureAlloc(g, kp.skipped, kp.skippedCapacity,
                        std::max<size_t>(size_t(kp.nSkipped), 1), "plan skipped samples"));
    kp.nItems = total;
    kp.nExtra = totals[1];
    kp.nWorkItems = totals[2];
    HWS_TRY(ensureAlloc(g, kp.itemBin, kp.itemBinCapacity, size_t(kp.nWorkItems),
                        "plan item bins"));
    const size_t occWords = (pl.nbins + 31) / 32;
    HWS_TRY(ensureAlloc(g, kp.occ, kp.occCapacity, occWords, "plan occupancy"));
    hipLaunchKernelGGL(occupancyKernel, dim3(blocksFor(occWords)), dim3(kThreads), 0, g->stream,
                       kp.binStart, pl.nbins, kp.occ);
    HWS_TRY(checkHip(g, hipGetLastError(), "occupancy launch"));
    const size_t nrowsBits = pl.nbins / size_t(pl.tilesX);
    const size_t rowWords = (nrowsBits + 31) / 32;
    HWS_TRY(ensureAlloc(g, kp.rowLive, kp.rowLiveCapacity, rowWords, "plan row bits"));
    hipLaunchKernelGGL(rowLiveKernel, dim3(blocksFor(nrowsBits)), dim3(kThreads), 0, g->stream,
                       kp.occ, nrowsBits, pl.tilesX, kp.rowLive);
    HWS_TRY(checkHip(g, hipGetLastError(), "row bits launch"));
    HWS_TRY(ensureAlloc(g, kp.need, kp.needCapacity, occWords, "plan need bits"));
    hipLaunchKernelGGL(needKernel, dim3(blocksFor(occWords)), dim3(kThreads), 0, g->stream,
                       kp.occ, pl.nbins, pl.tilesX, pl.tilesY, kp.need);
    HWS_TRY(checkHip(g, hipGetLastError(), "need launch"));
    hipLaunchKernelGGL(popcountKernel, dim3(blocksFor(occWords)), dim3(kThreads), 0, g->stream,
                       kp.occ, occWords, pl.counters + 8 + kind);
    hipLaunchKernelGGL(popcountKernel, dim3(blocksFor(occWords)), dim3(kThreads), 0, g->stream,
                       kp.need, occWords, pl.counters + 10 + kind);
    hipLaunchKernelGGL(popcountKernel, dim3(blocksFor(rowWords)), dim3(kThreads), 0, g->stream,
                       kp.rowLive, rowWords, pl.counters + 14 + kind);
    const size_t ncolsBits = pl.nbins / size_t(pl.tilesY);
    const size_t colWords = (ncolsBits + 31) / 32;
    HWS_TRY(ensureAlloc(g, kp.colNeed, kp.colNeedCapacity, colWords, "plan column bits"));
    hipLaunchKernelGGL(liveColumnsKernel, dim3(blocksFor(ncolsBits)), dim3(kThreads), 0,
                       g->stream, kp.need, ncolsBits, pl.tilesX, pl.tilesY, kp.colNeed);
    hipLaunchKernelGGL(popcountKernel, dim3(blocksFor(colWords)), dim3(kThreads), 0, g->stream,
                       kp.colNeed, colWords, pl.counters + 16 + kind);
    HWS_TRY(checkHip(g, hipGetLastError(), "tile count launch"));
    hipLaunchKernelGGL(itemBinsKernel, dim3(blocksFor(pl.nbins)), dim3(kThreads), 0, g->stream,
                       kp.itemStart, pl.nbins, kp.itemBin);
    HWS_TRY(checkHip(g, hipGetLastError(), "item bin launch"));
    HWS_TRY(ensureAlloc(g, kp.recSorted, kp.capacity, size_t(total), "plan sorted records"));
    if (kp.nExtra)
      HWS_TRY(ensureAlloc(g, kp.partials, kp.partialCapacity,
                          size_t(kp.nExtra) * kp.nImg * pl.tileSize * pl.tileSize,
                          "plan partial tiles"));
    // cursor = bin offsets
    HWS_TRY(checkHip(g, hipMemcpyAsync(kp.counts, kp.binStart, pl.nbins * sizeof(uint32_t),
                                       hipMemcpyDeviceToDevice, g->stream), "cursor copy"));
    ka.cursor = kp.counts;
    ka.recSorted = kp.recSorted;
    ka.nRecords = total;
    ka.skipped = kp.skipped;
    ka.nSkipped = uint32_t(kp.nSkipped);
    // sort keys bin << 12 | origin need 12 + log2(nbins) <= 32 bits
    int binBits = 0;
    while ((size_t(1) << binBits) < pl.nbins) ++binBits;
    // (the cell-owning grid kernel sums every record in registers, in any
    // order: it needs no sort)
    const bool sortByOrigin = binBits <= 20 && total > 1 && !cellKernelTile(pl.tileSize);
    if (sortByOrigin) {
      HWS_TRY(ensureAlloc(g, kp.keys, kp.keyCapacity, size_t(total), "plan sort keys"));
      HWS_TRY(ensureAlloc(g, kp.keysAlt, kp.keyAltCapacity, size_t(total), "plan sort keys"));
      HWS_TRY(ensureAlloc(g, kp.recAlt, kp.altCapacity, size_t(total), "plan sorted records"));
      ka.sortKeys = kp.keys;
    }
    if (n > 0) {
#define HWS_SCATTER(K, P)                                                                   \
  hipLaunchKernelGGL((binKernel<true, K, P>), dim3(sBlocks), dim3(kThreads), 0, g->stream,       \
                     dUvw, dWl, pl.rowOrder, nrow, int(nchan), sGroups, sStep, ga, ka)
      if (sortByOrigin)
        HWS_SCATTER(true, scatterPer);
      else
        HWS_SCATTER(false, scatterPer);
#undef HWS_SCATTER
      HWS_TRY(checkHip(g, hipGetLastError(), "bin scatter launch"));
    }
    hipLaunchKernelGGL(planCheckKernel, dim3(blocksFor(pl.nbins)), dim3(kThreads), 0, g->stream,
                       kp.counts, kp.binStart, pl.nbins, pl.counters + 12 + kind);
    HWS_TRY(checkHip(g, hipGetLastError(), "plan check launch"));
    if (sortByOrigin && n > 0) {
      uint32_t* k0 = kp.keys;
      uint2* r0 = kp.recSorted;
      HWS_TRY(sortRecords(g, kp.keys, kp.keysAlt, kp.recSorted, kp.recAlt, size_t(total),
                          12 + binBits));
      if (kp.keys != k0) std::swap(kp.keyCapacity, kp.keyAltCapacity);
      if (kp.recSorted != r0) std::swap(kp.capacity, kp.altCapacity);
    }
  }
  unsigned long long c[kPlanCounters] = {};
  HWS_TRY(checkHip(g, hipMemcpyAsync(c, pl.counters, sizeof(c), hipMemcpyDeviceToHost,
                                     g->stream), "plan check"));
  HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "plan check"));
  bool listed = true;
  for (int kind = 0; kind < 2; ++kind)
    if (pl.kinds[kind].active && c[18 + kind] != pl.kinds[kind].nSkipped) listed = false;
  if (c[12] || c[13] || !listed)
    return setLastError(g, HWS_ERR_BAD_STATE, "binning plan inconsistent: records misplaced");
  if (ga.rotate && n > 0) {
    HWS_TRY(ensureAlloc(g, pl.rot, pl.rotCapacity, n, "plan rotation phasors"));
    GridArgs ra{};
    ra.ga = ga;
    ra.nchan = nchan;
    hipLaunchKernelGGL(rotationTableKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                       g->stream, ra, dUvw, dWl, n, pl.rot);
    HWS_TRY(checkHip(g, hipGetLastError(), "rotation table launch"));
  }
  pl.valid = true;
  return HWS_OK;
}

namespace {
GridArgs gridArgs(const hws_gridder* g, int pass) {
  const Plan& pl = g->plan;
  GridArgs a;
  a.ga = geomArgs(g);
  a.nImgTotal = g->nImg;
  a.passStart = passLayerStart(g, pass);
  a.layerLo = a.passStart;
  a.layerHi = passLayerStart(g, pass + 1);
  a.nLayersPass = g->maxLayersPerPass;
  a.tilesPerLayer = pl.tilesX * pl.tilesY;
  a.sg = pl.tileSize + g->K - 1;
  a.nchan = pl.nchan;
  a.chunk = pl.chunk;
  a.rot = pl.rot;
  a.cw = (g->weights && pl.weighted) ? pl.cw : nullptr;
  return a;
}
}  // namespace

hws_status gridBinned(hws_gridder* g, const float* dVis, int pass, bool accumulate,
                      bool sparseLayers) {
  Plan& pl = g->plan;
  if (!pl.valid) return setLastError(g, HWS_ERR_BAD_STATE, "no binning plan");
  GridArgs a = gridArgs(g, pass);
  const int blocks = persistentBlocks(g, 8);
  const __half2* khalf = g->cfg.precision == HWS_PRECISION_MIXED ? g->dKernelPair : nullptr;
  const int tg = (g->K * g->K + 63) / 64;
  const float2* vis = reinterpret_cast<const float2*>(dVis);
  float2* layers = reinterpret_cast<float2*>(g->dLayers);
  const size_t ldsMax = maxLdsBytes(g);
  for (int kind = 0; kind < 2; ++kind) {
    const KindPlan& kp = pl.kinds[kind];
    if (!kp.active) continue;
    // images per launch: as many as the LDS tile allows
    int per = kp.nImg;
    while (per > 1 && gridLdsBytes(g, per, false) > ldsMax) --per;
    if (gridLdsBytes(g, per, false) > ldsMax)
      return setLastError(g, HWS_ERR_UNSUPPORTED, "uv tile and kernel table exceed LDS");
    for (int i0 = 0; i0 < kp.nImg; i0 += per) {
      const int n = std::min(per, kp.nImg - i0);
      a.imgs.n = n;
      for (int i = 0; i < kMaxImages; ++i) a.imgs.idx[i] = (i0 + i < kp.nImg) ? kp.img[i0 + i] : 0;
      const size_t lds = gridLdsBytes(g, n, false);
      if (!accumulate && !sparseLayers) {
        // the kernel writes the tiles of non-empty bins only: clear the pass's
        // layers of these images first (unless the caller reads them through
        // the occupancy bits)
        const size_t HW = size_t(g->W) * g->H;
        const size_t words = size_t(a.layerHi - a.layerLo) * HW * sizeof(float2) / 16;
        for (int i = 0; i < n; ++i) {
          uint4* base = reinterpret_cast<uint4*>(
              layers + (size_t(a.imgs.idx[i]) * a.nLayersPass + (a.layerLo - a.passStart)) * HW);
          const unsigned zb = unsigned(std::min<size_t>((words + kThreads - 1) / kThreads,
                                                        size_t(persistentBlocks(g, 16))));
          hipLaunchKernelGGL(zero16Kernel, dim3(std::max(1u, zb)), dim3(kThreads), 0, g->stream,
                             base, words);
        }
        HWS_TRY(checkHip(g, hipGetLastError(), "layer clear launch"));
      }
      // partial tiles of a launch hold its own images only
      if (mfmaKernelUsable(g) || mfmaF16(g)) {
        const unsigned mblocks = unsigned(std::max<uint64_t>(
            1, std::min<uint64_t>((kp.nWorkItems + kMfmaWaves - 1) / kMfmaWaves,
                                  uint64_t(persistentBlocks(g, 8)))));
        hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, g->stream, pl.queue, 0u, size_t(1));
        const bool f16 = mfmaF16(g), gt = f16 && f16TableGlobal(g);
        if (gt && !g->dF16Table)
          return setLastError(g, HWS_ERR_BAD_STATE, "padded kernel table not built");
        const float* gkt = g->dF16Table;
        const uint32_t* gkp =
            gt ? reinterpret_cast<const uint32_t*>(g->dF16Table + size_t(g->OS) *
                                                                      paddedRow(g->K, 16))
               : nullptr;
#define HWS_MFMA_GO(N, ACC, ROT)                                                               \
  if (gt)                                                                                      \
    hipLaunchKernelGGL((gridMfmaF16Kernel<N, ACC, ROT, true>), dim3(mblocks),                   \
                       dim3(kMfmaThreads), lds, g->stream, a, kp.binStart, kp.extraStart,       \
                       kp.itemStart, kp.itemBin, kp.recSorted, vis, g->dKernel, gkt, gkp,       \
                       layers, kp.partials, pl.queue, g->kernelExp);                            \
  else if (f16)                                                                                \
    hipLaunchKernelGGL((gridMfmaF16Kernel<N, ACC, ROT, false>), dim3(mblocks),                  \
                       dim3(kMfmaThreads), lds, g->stream, a, kp.binStart, kp.extraStart,       \
                       kp.itemStart, kp.itemBin, kp.recSorted, vis, g->dKernel, nullptr,        \
                       nullptr, layers, kp.partials, pl.queue, g->kernelExp);                   \
  else                                                                                         \
    hipLaunchKernelGGL((gridMfmaKernel<N, ACC, ROT>), dim3(mblocks), dim3(kMfmaThreads), lds,   \
                       g->stream, a, kp.binStart, kp.extraStart, kp.itemStart, kp.itemBin,      \
                       kp.recSorted, vis, g->dKernel, khalf, pl.uvw, pl.wavelengths, layers,    \
                       kp.partials, pl.queue)
#define HWS_MFMA_CASE(N)                                                                       \
  case N:                                                                                      \
    if (accumulate) {                                                                          \
      if (a.ga.rotate) {                                                                       \
        HWS_MFMA_GO(N, true, true);                                                            \
      } else {                                                                                 \
        HWS_MFMA_GO(N, true, false);                                                           \
      }                                                                                        \
    } else {                                                                                   \
      if (a.ga.rotate) {                                                                       \
        HWS_MFMA_GO(N, false, true);                                                           \
      } else {                                                                                 \
        HWS_MFMA_GO(N, false, false);                                                          \
      }                                                                                        \
    }                                                                                          \
    break;
        switch (n) {
          HWS_MFMA_CASE(1)
          HWS_MFMA_CASE(2)
          HWS_MFMA_CASE(3)
          HWS_MFMA_CASE(4)
        }
#undef HWS_MFMA_CASE
#undef HWS_MFMA_GO
      } else if (cellKernelTile(pl.tileSize)) {
        const bool pad = cellTablePadded(g->K, g->OS, pl.tileSize);
        const unsigned cblocks = unsigned(std::max<uint64_t>(
            1, std::min<uint64_t>((kp.nWorkItems + kCellWaves - 1) / kCellWaves,
                                  uint64_t(persistentBlocks(g, 8)))));
#define HWS_CELL_GO(N, TT, ACC, PAD, ROT)                                                   \
  hipLaunchKernelGGL((gridCellsKernel<N, TT, ACC, PAD, ROT>), dim3(cblocks), dim3(kCellThreads), \
                     lds, g->stream, a, kp.binStart, kp.extraStart, kp.itemStart, kp.itemBin,    \
                     kp.recSorted, vis, g->dKernel, khalf, pl.uvw, pl.wavelengths, layers,       \
                     kp.partials, pl.queue)
#define HWS_CELL_LAUNCH(N, TT, ACC)                \
  do {                                             \
    hipLaunchKernelGGL(fillKernel, dim3(1), dim3(64), 0, g->stream, pl.queue, 0u, size_t(1)); \
    if (pad && !a.ga.rotate)                       \
      HWS_CELL_GO(N, TT, ACC, true, false);        \
    else if (pad)                                  \
      HWS_CELL_GO(N, TT, ACC, true, true);         \
    else if (!a.ga.rotate)                         \
      HWS_CELL_GO(N, TT, ACC, false, false);       \
    else                                           \
      HWS_CELL_GO(N, TT, ACC, false, true);        \
  } while (0)
#define HWS_CELL_T(N, ACC)             \
  if (pl.tileSize == 16)               \
    HWS_CELL_LAUNCH(N, 16, ACC);       \
  else                                 \
    HWS_CELL_LAUNCH(N, 8, ACC);
#define HWS_CELL_CASE(N)               \
  case N:                              \
    if (accumulate) {                  \
      HWS_CELL_T(N, true)              \
    } else {                           \
      HWS_CELL_T(N, false)             \
    }                                  \
    break;
        switch (n) {
          HWS_CELL_CASE(1)
          HWS_CELL_CASE(2)
          HWS_CELL_CASE(3)
          HWS_CELL_CASE(4)
        }
#undef HWS_CELL_CASE
#undef HWS_CELL_T
#undef HWS_CELL_LAUNCH
#undef HWS_CELL_GO
      } else {
#define HWS_GRID_LAUNCH(N, TG, ACC)                                                           \
  hipLaunchKernelGGL((gridTilesKernel<N, TG, ACC>), dim3(blocks), dim3(kGridThreads), lds,      \
                     g->stream, a, kp.binStart, kp.extraStart, kp.itemStart, kp.itemBin,        \
                     kp.recSorted, vis,                                                         \
                     g->dKernel,                                                                \
                     khalf, pl.uvw, pl.wavelengths, layers, kp.partials)
#define HWS_GRID_TG(N, ACC)          \
  if (tg == 1)                       \
    HWS_GRID_LAUNCH(N, 1, ACC);      \
  else if (tg == 2)                  \
    HWS_GRID_LAUNCH(N, 2, ACC);      \
  else if (tg == 3)                  \
    HWS_GRID_LAUNCH(N, 3, ACC);      \
  else                               \
    HWS_GRID_LAUNCH(N, 4, ACC);
#define HWS_GRID_CASE(N)             \
  case N:                            \
    if (accumulate) {                \
      HWS_GRID_TG(N, true)           \
    } else {                         \
      HWS_GRID_TG(N, false)          \
    }                                \
    break;
      switch (n) {
        HWS_GRID_CASE(1)
        HWS_GRID_CASE(2)
        HWS_GRID_CASE(3)
        HWS_GRID_CASE(4)
      }
#undef HWS_GRID_CASE
#undef HWS_GRID_TG
#undef HWS_GRID_LAUNCH
      }
      HWS_TRY(checkHip(g, hipGetLastError(), "grid launch"));
      if (kp.nExtra) {
        const uint64_t cells = kThreads / kReduceParts;
        const uint64_t groups = (uint64_t(pl.tileSize) * pl.tileSize + cells - 1) / cells;
        const unsigned rblocks = unsigned(std::min<uint64_t>(kp.nExtra * groups, 65536));
        switch (n) {
#define HWS_REDUCE_CASE(N)                                                                    \
  case N:                                                                                     \
    hipLaunchKernelGGL(reducePartialsKernel<N>, dim3(rblocks), dim3(kThreads), 0, g->stream,   \
                       a, kp.extraStart, kp.partials, layers);                                \
    break;
          HWS_REDUCE_CASE(1)
          HWS_REDUCE_CASE(2)
          HWS_REDUCE_CASE(3)
          HWS_REDUCE_CASE(4)
#undef HWS_REDUCE_CASE
        }
        HWS_TRY(checkHip(g, hipGetLastError(), "partial reduce launch"));
      }
    }
  }
  return HWS_OK;
}

hws_status markSkipped(hws_gridder* g, float* dVisOut) {
  const Plan& pl = g->plan;
  if (!pl.valid) return setLastError(g, HWS_ERR_BAD_STATE, "no binning plan");
  for (int kind = 0; kind < 2; ++kind) {
    const KindPlan& kp = pl.kinds[kind];
    if (!kp.active || kp.nSkipped == 0) continue;
    ImgList imgs;
    imgs.n = kp.nImg;
    for (int i = 0; i < kMaxImages; ++i) imgs.idx[i] = kp.img[i];
    const size_t n = size_t(kp.nSkipped);
    hipLaunchKernelGGL(nanSkippedKernel, dim3(unsigned((n + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, g->stream, kp.skipped, n, imgs, g->nImg,
                       reinterpret_cast<float2*>(dVisOut));
    HWS_TRY(checkHip(g, hipGetLastError(), "not-predicted launch"));
  }
  return HWS_OK;
}

hws_status degridBinned(hws_gridder* g, float* dVisOut, int pass) {
  Plan& pl = g->plan;
  if (!pl.valid) return setLastError(g, HWS_ERR_BAD_STATE, "no binning plan");
  GridArgs a = gridArgs(g, pass);
  // as many workgroups as are resident at once (the kernel's registers allow
  // five waves per SIMD at two images): a static share of items per
  // workgroup, and no second round of workgroups running at low occupancy
  // (measured: 8 per CU, 13.2 ms per step at the 4096^2 benchmark; 5, 12.2)
  const int blocks = persistentBlocks(g, g->degridBlocksPerCu);
  const __half2* khalf = g->cfg.precision == HWS_PRECISION_MIXED ? g->dKernelPair : nullptr;
  const size_t ldsMax = maxLdsBytes(g);
  float2* out = reinterpret_cast<float2*>(dVisOut);
  for (int kind = 0; kind < 2; ++kind) {
    const KindPlan& kp = pl.kinds[kind];
    if (!kp.active) continue;
    // images per launch: as many subgrids as fit in LDS beside the kernel
    // table (each launch writes the samples of its own images only)
    int per = kp.nImg;
    while (per > 1 && gridLdsBytes(g, per, true) > ldsMax) --per;
    if (gridLdsBytes(g, per, true) > ldsMax)
      return setLastError(g, HWS_ERR_UNSUPPORTED, "uv subgrid and kernel table exceed LDS");
    for (int i0 = 0; i0 < kp.nImg; i0 += per) {
      const int n = std::min(per, kp.nImg - i0);
      a.imgs.n = n;
      for (int i = 0; i < kMaxImages; ++i)
        a.imgs.idx[i] = (i0 + i < kp.nImg) ? kp.img[i0 + i] : 0;
      const size_t lds = gridLdsBytes(g, n, true);
      switch (n) {
#define HWS_DEGRID_GO(N, KK)                                                                  \
  hipLaunchKernelGGL((degridTilesKernel<N, KK>), dim3(blocks), dim3(kThreads), lds, g->stream,   \
                     a, kp.binStart, kp.itemStart, kp.itemBin, kp.recSorted, g->dKernel, khalf,  \
                     pl.uvw, pl.wavelengths, reinterpret_cast<const float2*>(g->dLayers), out)
#define HWS_DEGRID_CASE(N)                                                                \
  case N:                                                                                   \
    if (g->K == 7)                                                                          \
      HWS_DEGRID_GO(N, 7);                                                                  \
    else                                                                                    \
      HWS_DEGRID_GO(N, 0);                                                                  \
    break;
        HWS_DEGRID_CASE(1)
        HWS_DEGRID_CASE(2)
        HWS_DEGRID_CASE(3)
        HWS_DEGRID_CASE(4)
#undef HWS_DEGRID_CASE
#undef HWS_DEGRID_GO
      }
      HWS_TRY(checkHip(g, hipGetLastError(), "degrid launch"));
    }
  }
  return HWS_OK;
}

namespace {
PlainArgs plainArgs(const hws_gridder* g, int img, int pass, size_t nrow,
                    size_t nchan) {
  PlainArgs a;
  a.W = g->W;
  a.H = g->H;
  a.K = g->K;
  a.OS = g->OS;
  a.mid = g->K / 2;
  a.dl = float(g->cfg.pixel_size_x);
  a.dm = float(g->cfg.pixel_size_y);
  a.nW = g->nW;
  a.minW = float(g->minW);
  a.maxW = float(g->maxW);
  a.isComplex = g->isComplex[img];
  a.passStart = passLayerStart(g, pass);
  a.layerLo = a.passStart;
  a.layerHi = passLayerStart(g, pass + 1);
  a.nLayersPass = g->maxLayersPerPass;
  a.img = img;
  a.nImgTotal = g->nImg;
  a.rotate = g->cfg.rotate_visibilities &&
             (g->cfg.l_shift != 0.0 || g->cfg.m_shift != 0.0);
  a.lShift = float(g->cfg.l_shift);
  a.mShift = float(g->cfg.m_shift);
  a.ch = 0;
  a.nrow = nrow;
  a.nchan = nchan;
  a.blocked = g->blocked;
  a.cw = (g->weights && g->plan.weighted) ? g->plan.cw : nullptr;
  return a;
}
}  // namespace

hws_status gridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl, const float* dVis,
                     int pass) {
  if (nrow == 0) return HWS_OK;
  for (int img = 0; img < g->nImg; ++img) {
    PlainArgs a = plainArgs(g, img, pass, nrow, nchan);
    for (size_t ch = 0; ch < nchan; ++ch) {
      a.ch = int(ch);
      hipLaunchKernelGGL(gridPlainKernel, dim3(blocksFor(nrow)), dim3(kThreads), 0, g->stream,
                         a, dUvw, dWl, reinterpret_cast<const float2*>(dVis), g->dKernel,
                         reinterpret_cast<float2*>(g->dLayers));
      HWS_TRY(checkHip(g, hipGetLastError(), "plain grid launch"));
    }
  }
  return HWS_OK;
}

hws_status degridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                       const double* dUvw, const double* dWl, float* dVisOut,
                       int pass) {
  if (nrow == 0) return HWS_OK;
  for (int img = 0; img < g->nImg; ++img) {
    PlainArgs a = plainArgs(g, img, pass, nrow, nchan);
    for (size_t ch = 0; ch < nchan; ++ch) {
      a.ch = int(ch);
      hipLaunchKernelGGL(degridPlainKernel, dim3(blocksFor(nrow)), dim3(kThreads), 0, g->stream,
                         a, dUvw, dWl, g->dKernel, reinterpret_cast<const float2*>(g->dLayers),
                         reinterpret_cast<float2*>(dVisOut));
      HWS_TRY(checkHip(g, hipGetLastError(), "plain degrid launch"));
    }
  }
  return HWS_OK;
}

namespace {
hws_status ensureNTable(hws_gridder* g, int order) {
  const size_t HW = size_t(g->W) * g->H;
  if (!g->dNTable)
    HWS_TRY(checkHip(g, hipMalloc(&g->dNTable, HW * sizeof(float)), "n' table"));
  if (g->nTableKind == order) return HWS_OK;
  hipLaunchKernelGGL(nTableKernel, dim3(blocksFor(HW)), dim3(kThreads), 0, g->stream, g->W, g->H,
                     int(g->blocked), g->cfg.pixel_size_x, g->cfg.pixel_size_y, g->cfg.l_shift, g->cfg.m_shift,
                     order, g->dNTable);
  HWS_TRY(checkHip(g, hipGetLastError(), "n' table launch"));
  g->nTableKind = order;
  return HWS_OK;
}
}  // namespace

hws_status projectPass(hws_gridder* g, int pass, bool fresh) {
  const size_t HW = size_t(g->W) * g->H;
  ImgArgs a = imgArgs(g, pass);
  a.fresh = fresh ? 1 : 0;
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
    HWS_TRY(ensureNTable(g, 0));
    for (int img = 0; img < g->nImg; ++img) {
      const int kind = g->isComplex[img];
      for (int li = 0; li < a.nLayersInPass; ++li) {
        const float w = float(layerToW(a.maps[kind], a.passStart + li));
        const float2* layer = reinterpret_cast<const float2*>(g->dLayers) +
                              (size_t(img) * a.nLayersPass + li) * HW;
        hipLaunchKernelGGL(projectPlainKernel, dim3(blocksFor(HW)), dim3(kThreads), 0, g->stream,
                           g->W, g->H, int(g->blocked), layer, g->dNTable, w, kind, g->dSum + img * HW,
                           g->dISum + img * HW);
        HWS_TRY(checkHip(g, hipGetLastError(), "plain project launch"));
      }
    }
    return HWS_OK;
  }
  hipLaunchKernelGGL(projectKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0, g->stream, a,
                     reinterpret_cast<const float2*>(g->dLayers), g->dSum, g->dComp, g->dISum,
                     g->dIComp);
  return checkHip(g, hipGetLastError(), "project launch");
}

hws_status projectFusedPass(hws_gridder* g, int pass, bool fresh) {
  const Plan& pl = g->plan;
  if (!pl.valid) return setLastError(g, HWS_ERR_BAD_STATE, "no binning plan");
  int tshift = 0;
  while ((1 << tshift) < pl.tileSize) ++tshift;
  const float2* layers = reinterpret_cast<const float2*>(g->dLayers);
  for (int kind = 0; kind < 2; ++kind) {
    const KindPlan& kp = pl.kinds[kind];
    if (!kp.active) continue;
    ColsProjectArgs a{};
    a.W = g->W;
    a.H = g->H;
    a.dl = g->cfg.pixel_size_x;
    a.dm = g->cfg.pixel_size_y;
    a.lShift = g->cfg.l_shift;
    a.mShift = g->cfg.m_shift;
    a.map = layerMap(g, kind);
    a.passStart = passLayerStart(g, pass);
    a.nLayersInPass = passLayerStart(g, pass + 1) - a.passStart;
    a.nLayersPass = g->maxLayersPerPass;
    a.fresh = fresh ? 1 : 0;
    a.rowLive = kp.rowLive;
    a.rowLiveWords = (pl.nbins / size_t(pl.tilesX) + 31) / 32;
    a.tilesY = pl.tilesY;
    a.tshift = tshift;
    a.nImg = kp.nImg;
    for (int i = 0; i < kMaxImages; ++i) a.img[i] = i < kp.nImg ? kp.img[i] : 0;
    HWS_TRY(launchColsProject(g, a, kind == 1, layers, g->dSum, g->dComp, g->dISum,
                              g->dIComp));
  }
  return HWS_OK;
}

hws_status finalizeImages(hws_gridder* g, double mult, float* dReal,
                          float* dImag) {
  const size_t HW = size_t(g->W) * g->H;
  const double norm = 1.0 / (double(g->OS) * double(g->OS));
  const int correct = g->cfg.kernel_mode != HWS_KERNEL_NEAREST_NEIGHBOUR;
  const double* wsum = (g->weights && g->weightSumsValid) ? g->dWeightSums : nullptr;
  hipLaunchKernelGGL(finalizeKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0, g->stream,
                     g->W, g->H, g->dSum, g->dComp, mult, g->dCorrXf, g->dCorrYf, norm, correct,
                     wsum, dReal);
  HWS_TRY(checkHip(g, hipGetLastError(), "finalize launch"));
  if (dImag && g->anyComplex) {
    hipLaunchKernelGGL(finalizeKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0, g->stream,
                       g->W, g->H, g->dISum, g->dIComp, mult, g->dCorrXf, g->dCorrYf, norm,
                       correct, wsum, dImag);
    HWS_TRY(checkHip(g, hipGetLastError(), "finalize launch"));
  }
  return HWS_OK;
}

hws_status initPrediction(hws_gridder* g, const float* dReal,
                          const float* dImag) {
  const size_t HW = size_t(g->W) * g->H;
  const double norm = 1.0 / (double(g->OS) * double(g->OS));
  const int correct = g->cfg.kernel_mode != HWS_KERNEL_NEAREST_NEIGHBOUR;
  hipLaunchKernelGGL(initPredictionKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0,
                     g->stream, g->W, g->H, g->cfg.pixel_size_x, g->cfg.pixel_size_y,
                     g->cfg.l_shift, g->cfg.m_shift, dReal, g->dModelRe, g->dCorrXf, g->dCorrYf,
                     norm, correct);
  HWS_TRY(checkHip(g, hipGetLastError(), "init prediction launch"));
  if (g->anyComplex && dImag) {
    hipLaunchKernelGGL(initPredictionKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0,
                       g->stream, g->W, g->H, g->cfg.pixel_size_x, g->cfg.pixel_size_y,
                       g->cfg.l_shift, g->cfg.m_shift, dImag, g->dModelIm, g->dCorrXf,
                       g->dCorrYf, norm, correct);
    HWS_TRY(checkHip(g, hipGetLastError(), "init prediction launch"));
  }
  return HWS_OK;
}

hws_status imageToLayers(hws_gridder* g, int pass) {
  const size_t HW = size_t(g->W) * g->H;
  const ImgArgs a = imgArgs(g, pass);
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
    HWS_TRY(ensureNTable(g, 1));
    for (int img = 0; img < g->nImg; ++img) {
      const int kind = g->isComplex[img];
      for (int li = 0; li < a.nLayersInPass; ++li) {
        const float w = float(layerToW(a.maps[kind], a.passStart + li));
        float2* layer = reinterpret_cast<float2*>(g->dLayers) + (size_t(img) * a.nLayersPass + li) * HW;
        hipLaunchKernelGGL(imageToLayerPlainKernel, dim3(blocksFor(HW)), dim3(kThreads), 0,
                           g->stream, g->W, g->H, int(g->blocked), g->dModelRe + img * HW,
                           g->dModelIm ? g->dModelIm + img * HW : nullptr, g->dNTable, w, kind,
                           layer);
        HWS_TRY(checkHip(g, hipGetLastError(), "plain image-to-layer launch"));
      }
    }
    return HWS_OK;
  }
  hipLaunchKernelGGL(imageToLayersKernel, dim3(blocksFor(HW), g->nImg), dim3(kThreads), 0,
                     g->stream, a, g->dModelRe, g->dModelIm,
                     reinterpret_cast<float2*>(g->dLayers));
  return checkHip(g, hipGetLastError(), "image-to-layers launch");
}

}  // namespace hws

// ------------------------------------------------------ imaging weights
namespace hws {
namespace {

constexpr int kWeightThreads = 256;

struct WeightSumArgs {
  LayerMap maps[2];
  int kindOf[kMaxImages];
  int mode;
  int viswVector;  // the visibility weights are aligned for vector loads
  WeightGridArgs wg;
};

// Per (row, channel) sample: the imaging weight GetWeight at (u, v) =
// uvw / lambda (gridding/msgridderbase.cpp:584-589, the float the reference
// keeps), and per image the weight the grid kernels multiply the visibility
// by, cw = visibilityWeight * imagingWeight (:567-589, the visibility weight
// after the weighting mode: as given, squared, or 1 where non-zero; a float
// product as the reference forms it). Per-block partial sums per image of
// the counters readAndWeightVisibilities keeps (:640-659), over the samples
// whose w-layer the image kind grids (the others are deselected, :563-565,
// and add nothing): [block][image][0 total weight, 1 visibility weight sum,
// 2 max gridded weight, 3 gridded count]. NIMG images per sample, read and
// written as one vector where NIMG is 2 or 4.
template <int NIMG>
__global__ void __launch_bounds__(kWeightThreads)
weightSumsKernel(const double* __restrict__ uvw, const double* __restrict__ wl, uint32_t nrow,
                 uint32_t nchan, const float* __restrict__ visw, WeightSumArgs a,
                 double* __restrict__ partials, float* __restrict__ cw) {
#pragma clang fp contract(off)
  __shared__ double red[kWeightThreads];
  double acc[NIMG][4];
#pragma unroll
  for (int i = 0; i < NIMG; ++i) acc[i][0] = acc[i][1] = acc[i][2] = acc[i][3] = 0.0;
  const uint32_t n = nrow * nchan;
  const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t s0 = blockIdx.x * per, s1 = s0 + per < n ? s0 + per : n;
  using Vec = typename std::conditional<NIMG == 4, float4,
                                        typename std::conditional<NIMG == 2, float2,
                                                                  float>::type>::type;
  constexpr bool kVec = NIMG == 2 || NIMG == 4;
  for (uint32_t s = s0 + threadIdx.x; s < s1; s += kWeightThreads) {
    const uint32_t row = s / nchan, ch = s - row * nchan;
    const double lambda = wl[ch];
    const double u = uvw[size_t(row) * 3] / lambda, v = uvw[size_t(row) * 3 + 1] / lambda;
    const double w = uvw[size_t(row) * 3 + 2] / lambda;
    const float imgw = float(imagingWeightAt(a.wg, u, v));
    const bool sel[2] = {layerOf(a.maps[0], w) >= 0, layerOf(a.maps[1], w) >= 0};
    float vw[NIMG], c[NIMG];
    if (visw) {
      if (kVec && a.viswVector) {
        const Vec x = reinterpret_cast<const Vec*>(visw)[s];
        const float* xf = reinterpret_cast<const float*>(&x);
#pragma unroll
        for (int i = 0; i < NIMG; ++i) vw[i] = xf[i];
      } else {
#pragma unroll
        for (int i = 0; i < NIMG; ++i) vw[i] = visw[size_t(s) * NIMG + i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < NIMG; ++i) vw[i] = 1.0f;
    }
#pragma unroll
    for (int i = 0; i < NIMG; ++i) {
      if (a.mode == 1)
        vw[i] = vw[i] * vw[i];
      else if (a.mode == 2)
        vw[i] = vw[i] != 0.0f ? 1.0f : 0.0f;
      c[i] = vw[i] * imgw;
    }
    if constexpr (kVec) {
      Vec y;
      float* yf = reinterpret_cast<float*>(&y);
#pragma unroll
      for (int i = 0; i < NIMG; ++i) yf[i] = c[i];
      reinterpret_cast<Vec*>(cw)[s] = y;
    } else {
#pragma unroll
      for (int i = 0; i < NIMG; ++i) cw[size_t(s) * NIMG + i] = c[i];
    }
#pragma unroll
    for (int i = 0; i < NIMG; ++i) {
      if (!sel[a.kindOf[i]]) continue;
      const double cum = double(c[i]);
      if (cum != 0.0) {
        acc[i][0] = acc[i][0] + cum;
        acc[i][1] = acc[i][1] + double(vw[i]);
        acc[i][2] = fmax(acc[i][2], cum);
        acc[i][3] += 1.0;
      }
    }
  }
  for (int i = 0; i < NIMG; ++i)
    for (int q = 0; q < 4; ++q) {
      red[threadIdx.x] = acc[i][q];
      __syncthreads();
      for (int off = kWeightThreads / 2; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) {
          const double o = red[threadIdx.x + off];
          red[threadIdx.x] = q == 2 ? fmax(red[threadIdx.x], o) : red[threadIdx.x] + o;
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) partials[(size_t(blockIdx.x) * kMaxImages + i) * 4 + q] = red[0];
      __syncthreads();
    }
}

// Sums the block partials in block order (one thread per image and counter).
__global__ void weightSumsFinalKernel(const double* __restrict__ partials, int nblocks, int nImg,
                                      double* __restrict__ sums) {
  const int t = threadIdx.x;
  if (t >= nImg * 4) return;
  const int i = t / 4, q = t - i * 4;
  double r = 0.0;
  for (int b = 0; b < nblocks; ++b) {
    const double v = partials[(size_t(b) * kMaxImages + i) * 4 + q];
    r = q == 2 ? fmax(r, v) : r + v;
  }
  sums[i * 4 + q] = r;
}

}  // namespace

hws_status planWeights(hws_gridder* g) {
  Plan& pl = g->plan;
  g->weightSumsValid = false;
  pl.weighted = false;
  if (!g->weights || !pl.valid) return HWS_OK;
  const size_t n = pl.nrow * pl.nchan;
  if (!g->dWeightSums)
    HWS_TRY(checkHip(g, hipMalloc(&g->dWeightSums, kMaxImages * 4 * sizeof(double)),
                     "weight sums"));
  const int nblocks = 1024;
  if (g->weightPartialCapacity < size_t(nblocks) * kMaxImages * 4) {
    if (g->dWeightPartials) (void)hipFree(g->dWeightPartials);
    g->dWeightPartials = nullptr;
    g->weightPartialCapacity = 0;
    HWS_TRY(checkHip(g, hipMalloc(&g->dWeightPartials,
                                  size_t(nblocks) * kMaxImages * 4 * sizeof(double)),
                     "weight partial sums"));
    g->weightPartialCapacity = size_t(nblocks) * kMaxImages * 4;
  }
  if (n > 0)
    HWS_TRY(ensureAlloc(g, pl.cw, pl.cwCapacity, n * size_t(g->nImg), "plan gridding weights"));
  WeightSumArgs a{};
  a.maps[0] = layerMap(g, 0);
  a.maps[1] = layerMap(g, 1);
  for (int i = 0; i < g->nImg; ++i) a.kindOf[i] = g->isComplex[i] ? 1 : 0;
  a.mode = g->visWeightMode;
  a.wg = weightGridArgs(g->weights);
  a.viswVector = (reinterpret_cast<uintptr_t>(g->dVisWeights) %
                  (sizeof(float) * size_t(g->nImg == 3 ? 1 : g->nImg))) == 0;
  const uint32_t nr = uint32_t(pl.nrow), nc = uint32_t(pl.nchan);
  switch (g->nImg) {
#define HWS_WSUMS(N)                                                                         \
  case N:                                                                                    \
    hipLaunchKernelGGL((weightSumsKernel<N>), dim3(nblocks), dim3(kWeightThreads), 0,         \
                       g->stream, pl.uvw, pl.wavelengths, nr, nc, g->dVisWeights, a,          \
                       g->dWeightPartials, pl.cw);                                           \
    break;
    HWS_WSUMS(1)
    HWS_WSUMS(2)
    HWS_WSUMS(3)
    HWS_WSUMS(4)
#undef HWS_WSUMS
    default:
      return setLastError(g, HWS_ERR_UNSUPPORTED, "images per context");
  }
  HWS_TRY(checkHip(g, hipGetLastError(), "weight sums launch"));
  hipLaunchKernelGGL(weightSumsFinalKernel, dim3(1), dim3(64), 0, g->stream,
                     g->dWeightPartials, nblocks, g->nImg, g->dWeightSums);
  HWS_TRY(checkHip(g, hipGetLastError(), "weight sums launch"));
  pl.weighted = true;
  g->weightSumsValid = true;
  return HWS_OK;
}

}  // namespace hws
