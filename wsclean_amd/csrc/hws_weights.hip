// Imaging weights on the GPU: ImageWeights (structures/imageweights.h:19-129,
// structures/imageweights.cpp:33-49, 75-144) for natural, uniform and Briggs
// weighting, gridded from device-resident rows.
//
// Grid: every (row, channel, polarisation) weight is added to the cell of its
// (u, v) in the v >= 0 half plane (rows with v < 0 are mirrored), and to the
// running total. One thread walks the channels of one row and merges the
// weights of consecutive channels that share a cell into one FP64 atomic
// add; the total is summed per block and the block sums in a fixed order.
// The reference adds weights one at a time in row order. Both orders give the
// same doubles whenever the sums are exact in FP64 (weights with at most
// about 53 - 24 - log2(samples per cell) significant bits between them, e.g.
// unit weights or the few-digit weights of a MeasurementSet); otherwise the
// grid agrees to FP64 rounding.
//
// FinishGridding: Briggs s^2 = (5 10^-R)^2 / (sum w^2 / sum w) and
// w -> 1 / (1 + w s^2); uniform w -> 1 / w; natural w -> 1 (cells with
// weight). The sum of squares is reduced in a fixed order; every product,
// sum and quotient is rounded as the reference rounds it (no contraction).
#include "hws_internal.h"

#include <cmath>
#include <string>
#include <vector>

#define HWS_TRY_NULL(expr)                   \
  do {                                       \
    const hws_status st_ = (expr);           \
    if (st_ != HWS_OK) return st_;           \
  } while (0)

namespace hws {
namespace {

constexpr int kWThreads = 256;

struct DeviceScope {
  int prev = -1;
  bool ok = false;
  explicit DeviceScope(int device) {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = hipSetDevice(device) == hipSuccess;
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

__device__ __forceinline__ double blockSum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = kWThreads / 2; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + off];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// ImageWeights::Grid(MSProvider&, BandData) (structures/imageweights.cpp:75-
// 111): one thread per row.
__global__ void __launch_bounds__(kWThreads)
weightGridKernel(const double* __restrict__ uvw, const double* __restrict__ wl, size_t nrow,
                 int nchan, int npol, const float* __restrict__ weights, int taper,
                 WeightGridArgs wg, double* __restrict__ grid, double* __restrict__ partials) {
  __shared__ double red[kWThreads];
  const size_t row = size_t(blockIdx.x) * kWThreads + threadIdx.x;
  double total = 0.0;
  if (row < nrow) {
    double uM = uvw[row * 3], vM = uvw[row * 3 + 1];
    if (vM < 0.0) {
      uM = -uM;
      vM = -vM;
    }
    long long cur = -1;
    double run = 0.0;
    for (int ch = 0; ch < nchan; ++ch) {
      const double u = uM / wl[ch], v = vM / wl[ch];
      const long long cell = weightCell(wg, u, v);
      if (cell < 0) continue;
      if (cell != cur) {
        if (cur >= 0) atomicAdd(grid + cur, run);
        cur = cell;
        run = 0.0;
      }
      const float* wp = weights ? weights + (row * size_t(nchan) + size_t(ch)) * size_t(npol)
                                : nullptr;
      for (int p = 0; p < npol; ++p) {
        float w = wp ? wp[p] : 1.0f;
        if (taper && w != 0.0f) w = 1.0f;
        run = run + double(w);
        total = total + double(w);
      }
    }
    if (cur >= 0) atomicAdd(grid + cur, run);
  }
  const double b = blockSum(total, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// Adds n partial sums to *total in a fixed order (one workgroup): thread t
// sums its contiguous share in index order, the shares are added by the
// fixed tree of blockSum, then the result to *total.
__global__ void __launch_bounds__(kWThreads)
addPartialsKernel(const double* __restrict__ partials, size_t n, double* __restrict__ total) {
#pragma clang fp contract(off)
  __shared__ double red[kWThreads];
  const size_t per = (n + kWThreads - 1) / kWThreads;
  const size_t i0 = size_t(threadIdx.x) * per, i1 = i0 + per < n ? i0 + per : n;
  double s = 0.0;
  for (size_t i = i0; i < i1; ++i) s = s + partials[i];
  const double b = blockSum(s, red);
  if (threadIdx.x == 0) *total = *total + b;
}

// Per-block partial sums of grid[i]^2 (FinishGridding's avgW numerator).
__global__ void __launch_bounds__(kWThreads)
sumSquaresKernel(const double* __restrict__ grid, size_t n, size_t perBlock,
                 double* __restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double red[kWThreads];
  const size_t b0 = size_t(blockIdx.x) * perBlock;
  const size_t b1 = b0 + perBlock < n ? b0 + perBlock : n;
  double s = 0.0;
  for (size_t i = b0 + threadIdx.x; i < b1; i += kWThreads)
    s = s + grid[i] * grid[i];
  const double b = blockSum(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

__global__ void __launch_bounds__(kWThreads)
finishKernel(double* __restrict__ grid, size_t n, int mode, double sSq) {
#pragma clang fp contract(off)
  const size_t i = size_t(blockIdx.x) * kWThreads + threadIdx.x;
  if (i >= n) return;
  const double v = grid[i];
  double r;
  if (mode == HWS_WEIGHT_BRIGGS)
    r = 1.0 / (1.0 + v * sSq);
  else if (mode == HWS_WEIGHT_UNIFORM)
    r = v != 0.0 ? 1.0 / v : 0.0;
  else
    r = v != 0.0 ? 1.0 : v;
  grid[i] = r;
}

hws_status fail(hws_status s, const std::string& msg) { return setLastError(nullptr, s, msg); }

hws_status ensurePartials(hws_weights* w, size_t n) {
  if (n <= w->partialCapacity) return HWS_OK;
  if (w->dPartials) (void)hipFree(w->dPartials);
  w->dPartials = nullptr;
  w->partialCapacity = 0;
  if (hipMalloc(&w->dPartials, n * sizeof(double)) != hipSuccess)
    return fail(HWS_ERR_OUT_OF_MEMORY, "weight grid partial sums");
  w->partialCapacity = n;
  return HWS_OK;
}

}  // namespace

WeightGridArgs weightGridArgs(const hws_weights* w) {
  WeightGridArgs a;
  a.grid = w->dGrid;
  a.width = w->width;
  a.height = w->height;
  a.dl = w->dl;
  a.dm = w->dm;
  return a;
}

}  // namespace hws

using namespace hws;

extern "C" hws_status hws_weights_create(const hws_weights_config* cfg, hws_weights** out) {
  if (!cfg || !out) return fail(HWS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (cfg->mode != HWS_WEIGHT_NATURAL && cfg->mode != HWS_WEIGHT_UNIFORM &&
      cfg->mode != HWS_WEIGHT_BRIGGS)
    return fail(HWS_ERR_INVALID_ARGUMENT, "unknown weighting mode");
  if (!(cfg->super_weight > 0.0) || !(cfg->pixel_scale_x > 0.0) || !(cfg->pixel_scale_y > 0.0))
    return fail(HWS_ERR_INVALID_ARGUMENT, "pixel scales and super weight must be positive");
  // the grid size as the constructor computes it (imageweights.cpp:33-49)
  size_t gw = size_t(std::round(double(cfg->width) / cfg->super_weight));
  size_t gh = size_t(std::round(double(cfg->height) / cfg->super_weight));
  if (gw % 2 != 0) ++gw;
  if (gh % 2 != 0) ++gh;
  if (gw == 0 || gh == 0 || gw > 65536 || gh > 65536)
    return fail(HWS_ERR_INVALID_ARGUMENT, "weight grid size out of range");
  DeviceScope scope(cfg->device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  hws_weights* w = new hws_weights();
  w->mode = cfg->mode;
  w->robust = cfg->briggs_robustness;
  w->weightsAsTaper = cfg->weights_as_taper;
  w->device = cfg->device;
  w->width = int(gw);
  w->height = int(gh);
  w->dl = cfg->pixel_scale_x;
  w->dm = cfg->pixel_scale_y;
  const size_t cells = gw * (gh / 2);
  if (hipMalloc(&w->dGrid, cells * sizeof(double)) != hipSuccess ||
      hipMalloc(&w->dTotal, sizeof(double)) != hipSuccess ||
      hipMemset(w->dGrid, 0, cells * sizeof(double)) != hipSuccess ||
      hipMemset(w->dTotal, 0, sizeof(double)) != hipSuccess) {
    hws_weights_destroy(w);
    return fail(HWS_ERR_OUT_OF_MEMORY, "weight grid allocation");
  }
  *out = w;
  return HWS_OK;
}

extern "C" void hws_weights_destroy(hws_weights* w) {
  if (!w) return;
  DeviceScope scope(w->device);
  if (w->dGrid) (void)hipFree(w->dGrid);
  if (w->dTotal) (void)hipFree(w->dTotal);
  if (w->dPartials) (void)hipFree(w->dPartials);
  delete w;
}

extern "C" hws_status hws_weights_grid_device(hws_weights* w, void* hip_stream, size_t nrow,
                                              size_t nchan, size_t npol,
                                              const double* d_uvw_m,
                                              const double* d_wavelengths,
                                              const float* d_weights) {
  if (!w) return fail(HWS_ERR_INVALID_ARGUMENT, "null weights");
  if (w->finished)
    return fail(HWS_ERR_BAD_STATE, "weights already finished (FinishGridding was called)");
  if (nrow == 0 || nchan == 0) return HWS_OK;
  if (npol == 0 || npol > 4 || nchan > (size_t(1) << 20) || nrow > (size_t(1) << 31))
    return fail(HWS_ERR_INVALID_ARGUMENT, "bad row, channel or polarisation count");
  if (!d_uvw_m || !d_wavelengths) return fail(HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceScope scope(w->device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  const size_t blocks = (nrow + kWThreads - 1) / kWThreads;
  HWS_TRY_NULL(ensurePartials(w, blocks));
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  hipLaunchKernelGGL(weightGridKernel, dim3(unsigned(blocks)), dim3(kWThreads), 0, stream,
                     d_uvw_m, d_wavelengths, nrow, int(nchan), int(npol), d_weights,
                     w->weightsAsTaper, weightGridArgs(w), w->dGrid, w->dPartials);
  hipLaunchKernelGGL(addPartialsKernel, dim3(1), dim3(kWThreads), 0, stream, w->dPartials, blocks,
                     w->dTotal);
  return checkHip(nullptr, hipGetLastError(), "weight gridding launch");
}

extern "C" hws_status hws_weights_finish(hws_weights* w, void* hip_stream) {
  if (!w) return fail(HWS_ERR_INVALID_ARGUMENT, "null weights");
  // imageweights.cpp:113-116
  if (w->finished) return fail(HWS_ERR_BAD_STATE, "FinishGridding() called twice");
  DeviceScope scope(w->device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  const size_t cells = size_t(w->width) * size_t(w->height / 2);
  double sSq = 0.0;
  if (w->mode == HWS_WEIGHT_BRIGGS) {
    const size_t perBlock = 4096;
    const size_t blocks = (cells + perBlock - 1) / perBlock;
    HWS_TRY_NULL(ensurePartials(w, blocks));
    hipLaunchKernelGGL(sumSquaresKernel, dim3(unsigned(blocks)), dim3(kWThreads), 0, stream,
                       w->dGrid, cells, perBlock, w->dPartials);
    HWS_TRY_NULL(checkHip(nullptr, hipGetLastError(), "weight sum of squares launch"));
    std::vector<double> part(blocks);
    double total = 0.0;
    HWS_TRY_NULL(checkHip(nullptr,
                          hipMemcpyAsync(part.data(), w->dPartials, blocks * sizeof(double),
                                         hipMemcpyDeviceToHost, stream),
                          "weight partial sums"));
    HWS_TRY_NULL(checkHip(nullptr,
                          hipMemcpyAsync(&total, w->dTotal, sizeof(double),
                                         hipMemcpyDeviceToHost, stream),
                          "weight total"));
    HWS_TRY_NULL(checkHip(nullptr, hipStreamSynchronize(stream), "weight finish"));
    double avgW = 0.0;
    for (double v : part) avgW = avgW + v;
    avgW = avgW / total;
    const double numeratorSqrt = 5.0 * std::exp((-M_LN10) * w->robust);
    sSq = numeratorSqrt * numeratorSqrt / avgW;
  }
  hipLaunchKernelGGL(finishKernel, dim3(unsigned((cells + kWThreads - 1) / kWThreads)),
                     dim3(kWThreads), 0, stream, w->dGrid, cells, w->mode, sSq);
  HWS_TRY_NULL(checkHip(nullptr, hipGetLastError(), "weight finish launch"));
  HWS_TRY_NULL(checkHip(nullptr, hipStreamSynchronize(stream), "weight finish"));
  w->finished = true;
  return HWS_OK;
}

extern "C" hws_status hws_weights_shape(const hws_weights* w, size_t* width, size_t* height) {
  if (!w || !width || !height) return fail(HWS_ERR_INVALID_ARGUMENT, "null argument");
  *width = size_t(w->width);
  *height = size_t(w->height);
  return HWS_OK;
}

extern "C" hws_status hws_weights_get_grid(const hws_weights* w, double* grid, double* total) {
  if (!w) return fail(HWS_ERR_INVALID_ARGUMENT, "null weights");
  DeviceScope scope(w->device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  HWS_TRY_NULL(checkHip(nullptr, hipDeviceSynchronize(), "weights"));
  const size_t cells = size_t(w->width) * size_t(w->height / 2);
  if (grid)
    HWS_TRY_NULL(checkHip(nullptr,
                          hipMemcpy(grid, w->dGrid, cells * sizeof(double),
                                    hipMemcpyDeviceToHost),
                          "weight grid download"));
  if (total)
    HWS_TRY_NULL(checkHip(nullptr, hipMemcpy(total, w->dTotal, sizeof(double),
                                             hipMemcpyDeviceToHost),
                          "weight total download"));
  return HWS_OK;
}
