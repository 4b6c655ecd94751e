// Internal declarations of the MI355X w-stacking library (not part of the
// public C ABI, which is include/hip_wstacking.h).
#ifndef HWS_INTERNAL_H
#define HWS_INTERNAL_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/hip_wstacking.h"

namespace hws {

constexpr int kMaxImages = 4;
constexpr int kPlanCounters = 20;
constexpr int kMaxFftStages = 24;
constexpr int kMaxKernelSize = 15;
constexpr int kMaxOversampling = 1023;

struct Error {
  hws_status status;
  std::string message;
};

// One 1-D transform length: the radices of its Stockham stages and the
// twiddle table exp(sign 2 pi i k / n) for both signs.
struct FftPlan {
  int n = 0;
  int nstages = 0;
  int radix[kMaxFftStages] = {};
  float2* twBackward = nullptr;  // sign +1 (FFTW_BACKWARD)
  float2* twForward = nullptr;   // sign -1 (FFTW_FORWARD)
};

struct FftPlanArgs {
  int n;
  int nstages;
  int radix[kMaxFftStages];
  const float2* tw;
};

// Layer map of one image kind (real images fold w < 0; complex images use
// [-maxW, maxW]), gridding/wstackinggridder.h:182-212.
struct LayerMap {
  int nW;
  double minW, maxW;
  int isComplex;
};

// The FP64 pixel-geometry helpers below round every product and sum on its
// own (no contraction into an FMA) or fuse explicitly, so that each kernel
// that evaluates them gets the same bits wherever the compiler inlines them.

// l and m of image pixel (x, y) (gridding/wstackinggridder.cpp:862-864,
// 922-924): l = (W/2 - x) dl + l_shift, m = (y - H/2) dm + m_shift.
__host__ __device__ inline double pixelL(int x, int W, double dl, double lShift) {
#pragma clang fp contract(off)
  return ((W / 2) - (double)x) * dl + lShift;
}
__host__ __device__ inline double pixelM(int y, int H, double dm, double mShift) {
#pragma clang fp contract(off)
  return ((double)y - (H / 2)) * dm + mShift;
}

// (l, m) inside the unit circle: l^2 + m^2 < 1.
__host__ __device__ inline bool nPrimeInside(double l, double m) {
#pragma clang fp contract(off)
  return l * l + m * m < 1.0;
}

// n' - 1 at (l, m) (0 outside the unit circle), as the reference's sqrt
// lookup tables hold it (gridding/wstackinggridder.cpp:851-872, 909-932).
__host__ __device__ inline double nPrime(double l, double m) {
#pragma clang fp contract(off)
  return (l * l + m * m < 1.0) ? sqrt(1.0 - l * l - m * m) - 1.0 : 0.0;
}

// The same from the power series of sqrt(1 - x) - 1 (x = l^2 + m^2) when
// x < 0.01, where the series to x^6 is exact to a few units of the last
// place of np (the next term is below 2e-16); the square root otherwise.
// Six FMAs instead of a correctly rounded FP64 square root, for the
// per-layer evaluation of the forward transform's input.
__host__ __device__ inline double nPrimeSeries(double l, double m) {
#pragma clang fp contract(off)
  const double x = l * l + m * m;
  if (!(x < 0.01)) return nPrime(l, m);
  // -x/2 - x^2/8 - x^3/16 - 5x^4/128 - 7x^5/256 - 21x^6/1024
  double p = -21.0 / 1024.0;
  p = fma(p, x, -7.0 / 256.0);
  p = fma(p, x, -5.0 / 128.0);
  p = fma(p, x, -1.0 / 16.0);
  p = fma(p, x, -1.0 / 8.0);
  p = fma(p, x, -1.0 / 2.0);
  return p * x;
}

#ifdef __HIPCC__
// sin and cos of 2 pi f for a fraction of a turn f in [-0.5, 0.5]: the
// hardware's v_sin_f32 and v_cos_f32, which take their argument in turns.
// Measured on 4M arguments spread over [-0.5, 0.5]: within 1.25e-7 of the
// exact values (RMS 3.5e-8), against 1e-7 for a Taylor evaluation after a
// quarter-turn reduction that costs 17 full-rate instructions where these
// are two quarter-rate ones.
__device__ __forceinline__ void turnSinCos(float f, float& s, float& c) {
  s = __builtin_amdgcn_sinf(f);
  c = __builtin_amdgcn_cosf(f);
}

__device__ __forceinline__ void kahanAdd(float& sum, float& comp, float term) {
  const float y = term - comp;
  const float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// One layer's term of projectOnImageAndCorrect (gridding/wstackinggridder.cpp:
// 874-907): the transformed layer value F at a pixel with n' = np, times
// exp(-2 pi i w n'). REF: FP64 sincos; otherwise the FP64-reduced fraction of
// a turn and turnSinCos. The products are fused explicitly and nothing else
// is (phi - rint(phi) would otherwise become an FMA in some kernels and not
// in others), so every kernel that projects gets the same bits.
template <bool REF>
__device__ __forceinline__ float2 projectTerm(float2 F, double w, double np) {
#pragma clang fp contract(off)
  if (REF) {
    double sn, cs;
    sincos(-2.0 * 3.14159265358979323846 * w * np, &sn, &cs);
    return make_float2(float(fma(double(F.x), cs, -(double(F.y) * sn))),
                       float(fma(double(F.x), sn, double(F.y) * cs)));
  }
  const double phi = w * np;
  const float f = float(phi - rint(phi));
  float sf, cf;
  turnSinCos(f, sf, cf);
  // exp(-2 pi i phi): sin = -sf, cos = cf
  return make_float2(fmaf(F.x, cf, F.y * sf), fmaf(-F.x, sf, F.y * cf));
}

// One layer cell of copyImageToLayerAndInverseCorrect (gridding/
// wstackinggridder.cpp:934-969): the model value (re, imv = -imaginary part
// for a complex image, else 0) at a pixel with n' = np, times
// exp(+2 pi i w n'). Precisions and explicit fusing as in projectTerm.
template <bool REF>
__device__ __forceinline__ float2 modelTerm(float re, float imv, bool cplx, double w, double np) {
#pragma clang fp contract(off)
  if (REF) {
    double sn, cs;
    sincos(2.0 * 3.14159265358979323846 * w * np, &sn, &cs);
    if (cplx)
      return make_float2(float(fma(double(re), cs, double(imv) * sn)),
                         float(fma(double(imv), cs, -(double(re) * sn))));
    return make_float2(float(double(re) * cs), float(-(double(re) * sn)));
  }
  const double phi = w * np;
  const float f = float(phi - rint(phi));
  float sn, cs;
  turnSinCos(f, sn, cs);
  if (cplx) return make_float2(fmaf(re, cs, imv * sn), fmaf(imv, cs, -(re * sn)));
  return make_float2(re * cs, -(re * sn));
}
#endif

// The imaging-weight grid of an ImageWeights (structures/imageweights.h:
// 73-94): `width` x `height`/2 doubles over the v >= 0 half of the uv plane.
struct WeightGridArgs {
  const double* grid;
  int width, height;
  double dl, dm;
};

#ifdef __HIPCC__
// Cell of (u, v) (wavelengths) in the weight grid, or -1 outside it:
// uvToXY + isWithinLimits (structures/imageweights.h:73-94). The products and
// the sum are rounded one by one in the reference's order, so a sample on a
// cell boundary lands in the reference's cell. The pragma keeps the compiler
// from contracting them into an FMA (HIP contracts by default, and __dmul_rn /
// __dadd_rn are plain operators in this toolchain's headers, which does not
// stop it).
__device__ inline long long weightCell(const WeightGridArgs& w, double u, double v) {
#pragma clang fp contract(off)
  if (v < 0.0) {
    u = -u;
    v = -v;
  }
  const double xf = u * double(w.width) * w.dl + double(w.width / 2);
  const double yf = v * double(w.height) * w.dm;
  const double fx = floor(xf), fy = floor(yf);
  // bounded before the integer conversion (the reference converts unchecked)
  if (!(fx >= 0.0) || !(fx < double(w.width)) || !(fy < double(w.height / 2)) || !(fy >= 0.0))
    return -1;
  return (long long)fx + (long long)fy * w.width;
}

// ImageWeights::GetWeight (structures/imageweights.h:34): the grid value at
// (u, v), 0 outside the grid.
__device__ inline double imagingWeightAt(const WeightGridArgs& w, double u, double v) {
  const long long c = weightCell(w, u, v);
  return c >= 0 ? w.grid[c] : 0.0;
}
#endif

__host__ __device__ inline double layerToW(const LayerMap& m, int layer) {
#pragma clang fp contract(off)
  if (m.nW == 1) return 0.0;
  if (m.isComplex) return layer * (m.maxW + m.maxW) / (m.nW - 1) - m.maxW;
  return layer * (m.maxW - m.minW) / (m.nW - 1) + m.minW;
}

// Binned samples of one image kind: one record per (sample, touched uv
// tile), sorted by (layer, tile) bin, with per-bin offsets.
struct KindPlan {
  bool active = false;
  int kind = 0;  // 0 real, 1 complex
  int nImg = 0;
  int img[kMaxImages] = {};
  uint64_t nItems = 0;
  uint2* recSorted = nullptr;    // records ordered by bin [capacity]
  uint32_t* counts = nullptr;    // histogram, then scatter cursor [nbins + 1]
  uint32_t* binStart = nullptr;  // exclusive scan of counts [nbins + 1]
  // A bin with more than `chunk` records is split into work items of at most
  // `chunk` records: item 0 writes the tile, items 1.. ("extra" items) leave
  // their tiles in `partials`, summed into the layer afterwards.
  // records in bin order are radix-sorted by (bin, footprint origin) so that
  // records sharing a footprint are consecutive: sort keys and the second
  // buffers of the sort
  uint32_t* keys = nullptr;
  uint32_t* keysAlt = nullptr;
  uint2* recAlt = nullptr;
  size_t keyCapacity = 0, keyAltCapacity = 0, altCapacity = 0;
  uint32_t* extraStart = nullptr;  // exclusive scan of extra items per bin [nbins + 1]
  uint32_t* itemStart = nullptr;   // exclusive scan of all items per bin [nbins + 1]
  uint64_t nWorkItems = 0;
  size_t itemCapacity = 0;
  uint32_t* itemBin = nullptr;  // bin of every work item [nWorkItems]
  size_t itemBinCapacity = 0;
  uint32_t* occ = nullptr;  // bit per bin: the bin holds records [(nbins + 31) / 32]
  size_t occCapacity = 0;
  // bit per bin: a degridding subgrid reads the tile (the bin or its left,
  // upper or upper-left neighbour holds records)
  uint32_t* need = nullptr;
  size_t needCapacity = 0;
  // bit per (layer, tile row): some tile of the row holds records
  uint32_t* rowLive = nullptr;
  size_t rowLiveCapacity = 0;
  // one bit per (layer, tile column) holding a tile the degridding reads
  uint32_t* colNeed = nullptr;
  size_t colNeedCapacity = 0;
  // samples of this kind that no record grids (off the grid or off every
  // layer): the predict writes NaN for them [nSkipped]
  uint32_t* skipped = nullptr;
  uint64_t nSkipped = 0;
  size_t skippedCapacity = 0;
  float2* partials = nullptr;      // [nExtra][nImg][T * T]
  uint64_t nExtra = 0;
  size_t extraCapacity = 0, partialCapacity = 0;
  size_t countsCapacity = 0, capacity = 0, binCapacity = 0;
};

struct Plan {
  bool valid = false;
  // made by hws_plan_device (the caller's device rows, which the planned
  // calls and graphs read) rather than by the host protocol's staged rows
  bool device = false;
  size_t nrow = 0, nchan = 0;
  const double* uvw = nullptr;         // device
  const double* wavelengths = nullptr;  // device
  int tilesX = 0, tilesY = 0, tileSize = 0;
  size_t nbins = 0;
  uint32_t chunk = 0;  // records per work item of the binned kernels
  KindPlan kinds[2];
  // temporaries
  uint32_t* scanTmp = nullptr;
  uint32_t* scanScratch = nullptr;
  size_t scanCapacity = 0, scanScratchCapacity = 0;
  // [kind * 3 + {0: gridded, 1: off grid, 2: off layer}], [6 + kind]: bins used
  // [0..2] gridded / off grid / off layer (real kind), [3..5] the same
  // (complex kind), [6 + kind] bins used, [8 + kind] occupied tiles,
  // [10 + kind] needed tiles, [12 + kind] bins failing the plan check,
  // [14 + kind] live (layer, tile row) pairs, [16 + kind] (layer, tile
  // column) pairs holding a needed tile, [18 + kind] samples listed as not
  // gridded
  unsigned long long* counters = nullptr;
  uint32_t* queue = nullptr;  // work-queue head of the running grid launch
  // rotating contexts: the gridding phase-centre phasor of every sample
  float2* rot = nullptr;
  size_t rotCapacity = 0;
  // rows sorted by their middle channel's bin: sort keys and row indices
  // (and the second buffers of the sort)
  uint32_t* rowKeys = nullptr;
  uint32_t* rowKeysAlt = nullptr;
  uint32_t* rowOrder = nullptr;
  uint32_t* rowOrderAlt = nullptr;
  size_t rowKeysCapacity = 0, rowKeysAltCapacity = 0, rowOrderCapacity = 0,
         rowOrderAltCapacity = 0;
  void* sortTemp = nullptr;  // radix sort scratch
  size_t sortTempCapacity = 0;
  // weighted contexts: the weight of every visibility, [sample][image]:
  // visibility weight (after the weighting mode) times the sample's imaging
  // weight (the float MSGridderBase keeps, gridding/msgridderbase.h:502), as
  // the grid kernels apply it
  float* cw = nullptr;
  size_t cwCapacity = 0;
  bool weighted = false;  // cw holds this plan's samples
};

// Version of what a captured graph of a context bakes in: the plan and its
// rows, the w-layer buffers and the imaging weights. Bumped whenever one of
// them changes; `alive` drops when the context is destroyed. Shared between
// the context and its graphs, so a graph outliving its context can tell.
struct PlanGeneration {
  std::atomic<uint64_t> value{0};
  std::atomic<bool> alive{true};
};

}  // namespace hws

// An ImageWeights on one device (structures/imageweights.{h,cpp}): the
// weight grid, the running total of gridded weights, and the scratch of the
// fixed-order reductions.
struct hws_weights {
  int mode = 0;  // hws_weight_mode
  double robust = 0.0;
  int weightsAsTaper = 0;
  int device = 0;
  int width = 0, height = 0;  // the grid holds width x height / 2 cells
  double dl = 0.0, dm = 0.0;
  double* dGrid = nullptr;
  double* dTotal = nullptr;     // sum of the gridded weights (one double)
  double* dPartials = nullptr;  // per-block partial sums
  size_t partialCapacity = 0;
  bool finished = false;
};

struct hws_gridder {
  hws_config cfg{};
  std::shared_ptr<hws::PlanGeneration> generation = std::make_shared<hws::PlanGeneration>();
  int W = 0, H = 0, K = 0, OS = 0;
  int nImg = 0;
  bool isComplex[hws::kMaxImages] = {};
  bool anyReal = false, anyComplex = false;
  hipStream_t ownStream = nullptr, stream = nullptr;
  bool userStream = false;  // stream was set by hws_set_stream (may be the null stream)
  bool capturing = false;   // a graph capture is in progress: no event records
  std::string err;

  // kernels
  std::vector<double> k1d;
  std::vector<double> corrX, corrY;  // D[k], k = 0..W/2 (H/2)
  float* dKernel = nullptr;          // [OS][K], tables[p][i]
  __half2* dKernelPair = nullptr;  // [OS][K] (hi, lo) FP16 pairs
  int kernelExp = 0;               // 2^kernelExp > the table's largest |value|
  // FP16 matrix-core grid kernel's padded table in device memory (FP32
  // values, then the (h, l) pairs), where it does not fit that kernel's LDS
  float* dF16Table = nullptr;
  float* dCorrXf = nullptr;  // float-rounded D tables (float reference)
  float* dCorrYf = nullptr;

  // layers
  int nW = 0, nPasses = 0, curPass = -1;
  double minW = 0, maxW = 0;
  int maxLayersPerPass = 0;
  float2* dLayers = nullptr;  // [nImg][maxLayersPerPass][H][W]
  bool layersZero = false;
  bool inversionActive = false, predictionActive = false;

  // image accumulators (Kahan sum + compensation) and model
  float* dSum = nullptr;   // [nImg][H][W]
  float* dComp = nullptr;
  float* dISum = nullptr;  // imaginary parts
  float* dIComp = nullptr;
  float* dModelRe = nullptr;
  float* dModelIm = nullptr;
  float* dNTable = nullptr;  // float n' table for the plain form
  int nTableKind = -1;       // 0 invert order, 1 predict order

  // FFT
  hws::FftPlan rowPlan, colPlan;
  // layer layout: blocked (four rows interleaved, see cellIndex) or natural
  bool blocked = false;
  // grid with the VALU cell kernel at T = 16 instead of the matrix-core one
  // (HWS_NO_MFMA=1 in the environment; for comparisons)
  bool noMfma = false;
  // the mixed precision grids with the FP32 matrix-core kernel
  // (HWS_NO_F16_MFMA=1)
  bool noF16Mfma = false;
  // degridding workgroups per CU (HWS_DEGRID_BLOCKS_PER_CU)
  int degridBlocksPerCu = 5;
  // lines (rows or columns) per workgroup of the blocked runtime-radix
  // transforms (HWS_FFT_LINES = 1, 2 or 4; 0: chosen by size)
  int fftLines = 0;
  // run the backward columns pass and the projection apart, and the forward
  // rows pass four layers per workgroup (the forms the fused kernels replace)
  bool noFusedProject = false;

  // binning plan
  hws::Plan plan;

  // host-API staging buffers
  double* dStageUvw = nullptr;
  double* dStageWl = nullptr;
  float* dStageVis = nullptr;
  float* dStageOut = nullptr;
  size_t stageRows = 0, stageVis = 0, stageWl = 0, stageOut = 0;

  // imaging weights (hws_set_weighting): the weight grid, the visibility
  // weights [sample][image] (null: 1) and their weighting mode, and the
  // per-image sums of the last plan: [image][0 total weight, 1 visibility
  // weight sum, 2 max gridded weight, 3 gridded count]
  const hws_weights* weights = nullptr;
  const float* dVisWeights = nullptr;
  int visWeightMode = 0;
  double* dWeightSums = nullptr;  // [kMaxImages][4]
  double* dWeightPartials = nullptr;
  size_t weightPartialCapacity = 0;
  bool weightSumsValid = false;

  // staged host rows (hws_stage_*): device copies of the rows' uvw and
  // visibilities, and two pinned host buffers the uploads go through
  double* dStagedUvw = nullptr;
  double* dStagedWl = nullptr;
  float* dStagedVis = nullptr;
  size_t stagedUvwCap = 0, stagedWlCap = 0, stagedVisCap = 0;
  unsigned char* pinned[2] = {nullptr, nullptr};
  hipEvent_t pinnedFree[2] = {nullptr, nullptr};
  int pinnedNext = 0;
  size_t stagedRows = 0, stagedMaxRows = 0, stagedNchan = 0;
  bool stagedWithVis = false, staging = false, staged = false;

  // timing
  hipEvent_t ev[8] = {};
  hws_stats stats{};
  int lastOp = 0;  // 1 invert, 2 predict (device one-call forms)
  int tileSize = 32;  // uv tile edge of the binned forms (divides W and H)
};

namespace hws {

// kernels / launchers (hws_kernels.hip)
// Message of the last failure of a call made without a context (per thread).
std::string& threadError();
// Records msg as g's last error (the thread's when g is null); returns s.
hws_status setLastError(hws_gridder* g, hws_status s, const std::string& msg);
hws_status checkHip(hws_gridder* g, hipError_t e, const char* what);

hws_status launchFill(hws_gridder* g, void* ptr, uint32_t value, size_t n32);
// Releases the staging buffers of hws_stage_* (hws_stage.hip).
void freeStaging(hws_gridder* g);
// The imaging weight of every planned sample and the per-image weight sums
// of the attached weighting (hws_set_weighting), on the context's stream.
hws_status planWeights(hws_gridder* g);
WeightGridArgs weightGridArgs(const hws_weights* w);
hws_status buildPlan(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl);
// sparseLayers: tiles without samples are left unwritten (not cleared); the
// caller reads the layers through the kinds' occupancy bits only.
hws_status gridBinned(hws_gridder* g, const float* dVis, int pass, bool accumulate,
                      bool sparseLayers = false);
hws_status degridBinned(hws_gridder* g, float* dVisOut, int pass);
// NaN for every sample of a kind that no record grids (with degridBinned
// over all passes, every value of the predicted visibilities is written).
hws_status markSkipped(hws_gridder* g, float* dVisOut);
// Exclusive scan of n 32-bit values into out[0..n] (out[n] = the total) on
// the context's stream; scratch holds scanScratchSize(n) values.
size_t scanScratchSize(size_t n);
hws_status exclusiveScan(hws_gridder* g, const uint32_t* in, uint32_t* out, size_t n,
                         uint32_t* scratch);
// Sorts n records (and their 32-bit keys) by key bits [0, endBit) on the
// context's stream (stable); on return `recs` holds the sorted records (the two record
// buffers may have been swapped).
hws_status sortRecords(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint2*& recs,
                       uint2*& recsAlt, size_t n, int endBit);
hws_status sortRows(hws_gridder* g, uint32_t*& keys, uint32_t*& keysAlt, uint32_t*& rows,
                    uint32_t*& rowsAlt, size_t n);
hws_status gridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl, const float* dVis,
                     int pass);
hws_status degridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                       const double* dUvw, const double* dWl, float* dVisOut,
                       int pass);
// fresh: the image accumulators start from zero instead of being read (the
// first pass of a device invert, which then needs no clear); ignored by the
// plain form.
hws_status projectPass(hws_gridder* g, int pass, bool fresh = false);
hws_status finalizeImages(hws_gridder* g, double mult, float* dReal,
                          float* dImag);
hws_status initPrediction(hws_gridder* g, const float* dReal,
                          const float* dImag);
hws_status imageToLayers(hws_gridder* g, int pass);
// The projection of a pass fused with its backward columns pass (after the
// sparse rows pass): both image kinds.
hws_status projectFusedPass(hws_gridder* g, int pass, bool fresh);

// FFT (hws_fft.hip)
hws_status makeFftPlan(hws_gridder* g, int n, FftPlan& plan);
void freeFftPlan(FftPlan& plan);
// Unnormalised 2-D transform of nbatch H x W complex arrays, in place.
// Sparse layer transforms (blocked layout only). Backward: uv tiles without
// samples are skipped on input (not read, need not be cleared; occ = the
// occupancy bits). Forward: only the tiles a degridding subgrid reads are
// written on output (occ = the need bits). See fftRowsBlockedSparseKernel and
// fftColsBlockedSparseKernel.
struct FftSparse {
  const uint32_t* rowLive;  // backward: one bit per (layer, tile row) holding samples
  const uint32_t* occ;  // one bit per (layer, tile) bin
  int layer0;           // layer of the first transform of the batch
  int tilesY;
  int tshift;           // log2 of the tile edge
  bool rowsOnly = false;  // backward: the columns pass runs fused with the projection
  // forward from a model: one bit per (layer, tile column) holding a needed
  // tile (the row pass then writes those columns only), or null
  const uint32_t* colNeed = nullptr;
};

// The columns pass of a pass's backward transforms fused with the projection
// of its layers onto the image accumulators (fftColsProjectKernel): the
// images of one kind, after their sparse backward rows pass.
struct ColsProjectArgs {
  int W, H;
  double dl, dm, lShift, mShift;
  LayerMap map;
  int passStart, nLayersInPass, nLayersPass;
  int fresh;  // the accumulators start from zero (not read)
  const uint32_t* rowLive;  // one bit per (layer, tile row) holding samples
  size_t rowLiveWords;
  int tilesY, tshift;
  int nImg;
  int img[kMaxImages];
};
// Whether the fused columns pass and projection covers a transform of
// height H with this many layers in a pass.
bool colsProjectSupported(const hws_gridder* g, int nLayersInPass);
hws_status launchColsProject(hws_gridder* g, const ColsProjectArgs& a, bool cplx,
                             const float2* layers, float* sum, float* comp, float* isum,
                             float* icomp);
// The model image a forward transform's row pass reads instead of the
// layers: each layer cell is computed on the fly as imageToLayersKernel
// writes it (copyImageToLayerAndInverseCorrect, wstackinggridder.cpp:934-969).
struct ModelSource {
  const float* re;  // this image's model (remapped and corrected), row-major
  const float* im;  // imaginary part (complex images) or null
  int cplx;
  int precision;
  double dl, dm, lShift, mShift;
  LayerMap map;
  int layer0;  // absolute layer of the first transform of the batch
};
hws_status fft2d(hws_gridder* g, float2* data, int nbatch, int sign,
                 const FftSparse* sparse = nullptr, const ModelSource* model = nullptr);

// Storage index of layer cell (x, y): natural row-major, or blocked with
// groups of four rows interleaved element by element.
__host__ __device__ inline size_t cellIndex(int x, int y, int W, bool blocked) {
  return blocked ? ((size_t(y >> 2) * size_t(W) + size_t(x)) << 2) + size_t(y & 3)
                 : size_t(y) * size_t(W) + size_t(x);
}

// Inverse of cellIndex for storage index p.
__host__ __device__ inline void cellOf(size_t p, int W, bool blocked, int& x, int& y) {
  if (blocked) {
    const size_t rowBlock = p / (size_t(4) * W);
    const int rem = int(p - rowBlock * size_t(4) * W);
    x = rem >> 2;
    y = int(rowBlock) * 4 + (rem & 3);
  } else {
    y = int(p / size_t(W));
    x = int(p - size_t(y) * W);
  }
}

bool blockedLayoutSupported(int W, int H);
// Builds dF16Table where the mixed precision's FP16 grid kernel reads its
// padded table from device memory (no-op otherwise).
hws_status buildF16Table(hws_gridder* g);
// Whether the runtime-radix transforms of the blocked layout fit one
// workgroup at this gridder's sizes (needs the row and column plans).
bool blockedGenericFits(const hws_gridder* g);
bool fastBlockedSize(int n);

inline int passLayerStart(const hws_gridder* g, int pass) {
  return int((size_t(g->nW) * size_t(pass)) / size_t(g->nPasses));
}

LayerMap layerMap(const hws_gridder* g, int kind);

}  // namespace hws

#endif
