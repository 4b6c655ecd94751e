// Internal declarations of the MI355X w-stacking library (not part of the
// public C ABI, which is include/hip_wstacking.h).
#ifndef HWS_INTERNAL_H
#define HWS_INTERNAL_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hip_wstacking.h"

namespace hws {

constexpr int kMaxImages = 4;
constexpr int kMaxFftStages = 24;
constexpr int kMaxKernelSize = 15;
constexpr int kMaxOversampling = 1023;
constexpr uint32_t kChunkItems = 1024;  // max samples per binned work item

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

__host__ __device__ inline double layerToW(const LayerMap& m, int layer) {
  if (m.nW == 1) return 0.0;
  if (m.isComplex) return layer * (m.maxW + m.maxW) / (m.nW - 1) - m.maxW;
  return layer * (m.maxW - m.minW) / (m.nW - 1) + m.minW;
}

// Binned samples of one image kind: sorted records + per-bin offsets +
// chunk list (bin, first record, count).
struct KindPlan {
  bool active = false;
  int kind = 0;  // 0 real, 1 complex
  int nImg = 0;
  int img[kMaxImages] = {};
  uint64_t nItems = 0;
  uint2* recSorted = nullptr;   // [capacity]
  uint32_t* binStart = nullptr;  // [nbins + 1]
  uint32_t* chunkStart = nullptr;  // [nbins + 1]
  uint4* chunks = nullptr;         // [capacity chunks]
  size_t chunkCapacity = 0;
  size_t chunkStartCapacity = 0;
  size_t capacity = 0;
  size_t binCapacity = 0;
};

struct Plan {
  bool valid = false;
  size_t nrow = 0, nchan = 0;
  const double* uvw = nullptr;         // device
  const double* wavelengths = nullptr;  // device
  int tilesX = 0, tilesY = 0, tileSize = 0;
  size_t nbins = 0;
  KindPlan kinds[2];
  // temporaries
  uint2* recTmp = nullptr;
  uint32_t* binTmp = nullptr;
  uint32_t* counts = nullptr;
  uint32_t* scanTmp = nullptr;
  uint32_t* scanScratch = nullptr;
  size_t tmpCapacity = 0, binTmpCapacity = 0, countsCapacity = 0, scanCapacity = 0;
  size_t scanScratchCapacity = 0;
  unsigned long long* counters = nullptr;  // [4]: gridded, off grid, off layer, items
};

}  // namespace hws

struct hws_gridder {
  hws_config cfg{};
  int W = 0, H = 0, K = 0, OS = 0;
  int nImg = 0;
  bool isComplex[hws::kMaxImages] = {};
  bool anyReal = false, anyComplex = false;
  hipStream_t ownStream = nullptr, stream = nullptr;
  std::string err;

  // kernels
  std::vector<double> k1d;
  std::vector<double> corrX, corrY;  // D[k], k = 0..W/2 (H/2)
  float* dKernel = nullptr;          // [OS][K], tables[p][i]
  __half2* dKernelPair = nullptr;  // [OS][K] (hi, lo) FP16 pairs
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

  // binning plan
  hws::Plan plan;

  // host-API staging buffers
  double* dStageUvw = nullptr;
  double* dStageWl = nullptr;
  float* dStageVis = nullptr;
  float* dStageOut = nullptr;
  size_t stageRows = 0, stageVis = 0, stageWl = 0, stageOut = 0;

  // timing
  hipEvent_t ev[8] = {};
  hws_stats stats{};
  int lastOp = 0;  // 1 invert, 2 predict (device one-call forms)
};

namespace hws {

// kernels / launchers (hws_kernels.hip)
hws_status setLastError(hws_gridder* g, hws_status s, const std::string& msg);
hws_status checkHip(hws_gridder* g, hipError_t e, const char* what);

hws_status launchFill(hws_gridder* g, void* ptr, uint32_t value, size_t n32);
hws_status buildPlan(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl);
hws_status gridBinned(hws_gridder* g, const float* dVis, int pass);
hws_status degridBinned(hws_gridder* g, float* dVisOut, int pass);
hws_status gridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* dUvw, const double* dWl, const float* dVis,
                     int pass);
hws_status degridPlain(hws_gridder* g, size_t nrow, size_t nchan,
                       const double* dUvw, const double* dWl, float* dVisOut,
                       int pass);
hws_status projectPass(hws_gridder* g, int pass);
hws_status finalizeImages(hws_gridder* g, double mult, float* dReal,
                          float* dImag);
hws_status initPrediction(hws_gridder* g, const float* dReal,
                          const float* dImag);
hws_status imageToLayers(hws_gridder* g, int pass);

// FFT (hws_fft.hip)
hws_status makeFftPlan(hws_gridder* g, int n, FftPlan& plan);
void freeFftPlan(FftPlan& plan);
// Unnormalised 2-D transform of nbatch H x W complex arrays, in place.
hws_status fft2d(hws_gridder* g, float2* data, int nbatch, int sign);

inline int passLayerStart(const hws_gridder* g, int pass) {
  return int((size_t(g->nW) * size_t(pass)) / size_t(g->nPasses));
}

LayerMap layerMap(const hws_gridder* g, int kind);

}  // namespace hws

#endif
