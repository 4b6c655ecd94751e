// C ABI of the MI355X w-stacking core (include/hip_wstacking.h): context
// set-up, kernel tables, w-layer planning and the pass protocol of
// WStackingGridder<T> (gridding/wstackinggridder.h:44-84).
#include "hws_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <cstdlib>
#include <new>
#include <vector>

using namespace hws;

namespace hws {
std::string& threadError() {
  thread_local std::string msg;
  return msg;
}
}  // namespace hws

namespace {

#define HWS_TRY(expr)            \
  do {                           \
    hws_status _s = (expr);      \
    if (_s != HWS_OK) return _s; \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

// hws_add_data / hws_sample_data plan each block of rows they are given,
// which replaces a plan made by hws_plan_device.
constexpr const char* kReplacedPlan =
    "the device plan was replaced by hws_add_data or hws_sample_data: call hws_plan_device "
    "again";

// ---- 1-D kernels: the closed forms of gridding/wstackinggridder.cpp:425-515
double bessel0(double x, double precision) {
  double d = 0.0, ds = 1.0, sum = 1.0;
  do {
    d += 2.0;
    ds *= x * x / (d * d);
    sum += ds;
  } while (ds > sum * precision);
  return sum;
}

void kaiserBessel(std::vector<double>& k, double alpha, size_t os, bool withSinc) {
  const size_t n = k.size(), mid = n / 2;
  const double ratio = 1.0 / double(os);
  const double norm = double(os) / bessel0(alpha, 1e-8);
  for (size_t i = 0; i != mid + 1; ++i) {
    const double sinc = (i == 0) ? ratio
                        : withSinc ? std::sin(kPi * ratio * double(i)) / (kPi * double(i))
                                   : ratio;
    const double t = double(i) / mid;
    k[mid + i] = sinc * bessel0(alpha * std::sqrt(1.0 - t * t), 1e-8) * norm;
  }
  for (size_t i = 0; i != mid; ++i) k[i] = k[n - 1 - i];
}

void rectangular(std::vector<double>& k, size_t os) {
  const size_t n = k.size(), mid = n / 2;
  const double ratio = 1.0 / double(os);
  k[mid] = 1.0;
  for (size_t i = 1; i != mid + 1; ++i)
    k[mid + i] = double(os) * std::sin(kPi * ratio * double(i)) / (kPi * double(i));
  for (size_t i = 0; i != mid; ++i) k[i] = k[n - 1 - i];
}

void gaussian(std::vector<double>& k, size_t os, bool withSinc) {
  const size_t n = k.size(), mid = n / 2;
  const double ratio = 1.0 / double(os);
  k[mid] = 1.0;
  for (size_t i = 1; i != mid + 1; ++i) {
    const double x = double(i), y = x * x * 3.0 * 3.0 / double(mid * mid);
    k[mid + i] = withSinc ? double(os) * std::sin(kPi * ratio * x) / (kPi * x) * std::exp(y / -2.0)
                          : std::exp(y / -2.0);
  }
  for (size_t i = 0; i != mid; ++i) k[i] = k[n - 1 - i];
}

void blackmanNuttall(std::vector<double>& k, size_t os, bool withSinc) {
  const size_t n = k.size(), mid = n / 2;
  const double ratio = 1.0 / double(os);
  const double a0 = 0.3635819, a1 = 0.4891775, a2 = 0.1365995, a3 = 0.0106411;
  k[mid] = 1.0;
  for (size_t i = 1; i != mid + 1; ++i) {
    const double x = double(i), y = double(mid + i);
    const double bn = a0 - a1 * std::cos(2.0 * kPi * y / double(n - 1)) +
                      a2 * std::cos(4.0 * kPi * y / double(n - 1)) -
                      a3 * std::cos(6.0 * kPi * y / double(n - 1));
    k[mid + i] = withSinc ? double(os) * std::sin(kPi * ratio * x) / (kPi * x) * bn : bn;
  }
  for (size_t i = 0; i != mid; ++i) k[i] = k[n - 1 - i];
}

// Builds the 1-D kernel of `mode`. nnIsDelta selects GetKernel's delta for
// nearest neighbour; the gridder itself keeps the KB kernel in that mode.
bool makeKernel(int mode, size_t os, size_t size, bool nnIsDelta,
                std::vector<double>& k) {
  k.assign(os * size, 0.0);
  const double alpha = 8.6;
  switch (mode) {
    case HWS_KERNEL_NEAREST_NEIGHBOUR:
      if (nnIsDelta)
        k[os * size / 2] = 1.0;
      else
        kaiserBessel(k, alpha, os, true);
      return true;
    case HWS_KERNEL_KAISER_BESSEL: kaiserBessel(k, alpha, os, true); return true;
    case HWS_KERNEL_KAISER_BESSEL_NO_SINC: kaiserBessel(k, alpha, os, false); return true;
    case HWS_KERNEL_RECTANGULAR: rectangular(k, os); return true;
    case HWS_KERNEL_GAUSSIAN: gaussian(k, os, true); return true;
    case HWS_KERNEL_GAUSSIAN_NO_SINC: gaussian(k, os, false); return true;
    case HWS_KERNEL_BLACKMAN_NUTTALL: blackmanNuttall(k, os, true); return true;
    case HWS_KERNEL_BLACKMAN_NUTTALL_NO_SINC: blackmanNuttall(k, os, false); return true;
    default: return false;
  }
}

// REDFT01 of the half kernel (gridding/wstackinggridder.cpp:784-789) at
// k = 0..axis/2: D[k] = h0 + 2 sum_{j>=1} h_j cos(pi j (k + 1/2) / n), n = axis*OS/2.
std::vector<double> kernelCorrection(const std::vector<double>& k1d, size_t K,
                                     size_t OS, size_t axis) {
  const size_t n = axis * OS / 2, half = K * OS / 2;
  std::vector<double> out(axis / 2 + 1);
  for (size_t k = 0; k != out.size(); ++k) {
    long double acc = k1d[half];
    for (size_t j = 1; j <= half && j < n; ++j)
      acc += 2.0L * (long double)k1d[half + j] *
             cosl(3.141592653589793238462643383279502884L * (long double)j *
                  ((long double)k + 0.5L) / (long double)n);
    out[k] = double(acc);
  }
  return out;
}

bool validSize(size_t n) {
  if (n == 0 || n % 2 != 0 || n > 16384) return false;
  size_t m = n;
  for (size_t p : {2, 3, 5, 7})
    while (m % p == 0) m /= p;
  return m == 1;
}

hws_status fail(hws_gridder* g, hws_status s, const std::string& msg) {
  if (g)
    g->err = msg;
  else
    threadError() = msg;
  return s;
}

template <typename T>
void freeDev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
hws_status devAlloc(hws_gridder* g, T*& p, size_t count, const char* what) {
  freeDev(p);
  return checkHip(g, hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), what);
}

template <typename T>
hws_status stageAlloc(hws_gridder* g, T*& p, size_t& cap, size_t count,
                      const char* what) {
  if (p && cap >= count) return HWS_OK;
  freeDev(p);
  cap = 0;
  HWS_TRY(checkHip(g, hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), what));
  cap = count;
  return HWS_OK;
}

size_t layersBytes(const hws_gridder* g, size_t layers) {
  return size_t(g->nImg) * layers * size_t(g->W) * size_t(g->H) * sizeof(float2);
}

hws_status zeroLayers(hws_gridder* g) {
  if (g->layersZero) return HWS_OK;
  HWS_TRY(launchFill(g, g->dLayers, 0u, layersBytes(g, g->maxLayersPerPass) / 4));
  g->layersZero = true;
  return HWS_OK;
}

hws_status zeroAccumulators(hws_gridder* g) {
  const size_t n = size_t(g->nImg) * g->W * g->H;
  HWS_TRY(launchFill(g, g->dSum, 0u, n));
  HWS_TRY(launchFill(g, g->dComp, 0u, n));
  if (g->anyComplex) {
    HWS_TRY(launchFill(g, g->dISum, 0u, n));
    HWS_TRY(launchFill(g, g->dIComp, 0u, n));
  }
  return HWS_OK;
}

int layersInPass(const hws_gridder* g, int pass) {
  return passLayerStart(g, pass + 1) - passLayerStart(g, pass);
}

// sparse, backward: the pass's layers hold only the tiles of bins with
// records (the binned grid without the clear) and the transform skips the
// others. sparse, forward: only the tiles the binned degridding reads are
// written.
// fromModel (forward, with sparse): the row pass computes the layers from the
// model images (imageToLayers is not run).
hws_status fftPassLayers(hws_gridder* g, int pass, int sign, bool sparse = false,
                         bool fromModel = false, bool rowsOnly = false) {
  const int n = layersInPass(g, pass);
  const size_t HW = size_t(g->W) * g->H;
  for (int img = 0; img < g->nImg; ++img) {
    ModelSource ms{};
    if (fromModel) {
      ms.re = g->dModelRe + size_t(img) * HW;
      ms.im = g->dModelIm ? g->dModelIm + size_t(img) * HW : nullptr;
      ms.cplx = g->isComplex[img] ? 1 : 0;
      ms.precision = g->cfg.precision;
      ms.dl = g->cfg.pixel_size_x;
      ms.dm = g->cfg.pixel_size_y;
      ms.lShift = g->cfg.l_shift;
      ms.mShift = g->cfg.m_shift;
      ms.map = layerMap(g, ms.cplx);
      ms.layer0 = passLayerStart(g, pass);
    }
    FftSparse sp{};
    if (sparse) {
      const Plan& pl = g->plan;
      const KindPlan& kp = pl.kinds[g->isComplex[img] ? 1 : 0];
      sp.occ = sign > 0 ? kp.occ : kp.need;
      sp.rowLive = kp.rowLive;
      sp.colNeed = (sign < 0 && fromModel && !g->noFusedProject) ? kp.colNeed : nullptr;
      sp.layer0 = passLayerStart(g, pass);
      sp.tilesY = pl.tilesY;
      sp.tshift = 0;
      while ((1 << sp.tshift) < pl.tileSize) ++sp.tshift;
      sp.rowsOnly = rowsOnly;
    }
    HWS_TRY(fft2d(g, g->dLayers + size_t(img) * g->maxLayersPerPass * HW, n, sign,
                  sparse ? &sp : nullptr, fromModel ? &ms : nullptr));
  }
  return HWS_OK;
}

// The binned invert leaves empty tiles unwritten when the blocked transform
// can skip them (power-of-two tile edge of at least 4).
bool sparseInvert(const hws_gridder* g) {
  const int T = g->plan.tileSize;
  return g->blocked && g->cfg.precision != HWS_PRECISION_PLAIN_FP32 && T >= 4 &&
         (T & (T - 1)) == 0;
}

bool prepared(const hws_gridder* g) { return g->nPasses > 0 && g->dLayers; }

hws_status stageRows(hws_gridder* g, size_t nrow, size_t nchan,
                     const double* uvw, const double* wl, size_t visFloats) {
  HWS_TRY(stageAlloc(g, g->dStageUvw, g->stageRows, nrow * 3, "stage uvw"));
  HWS_TRY(stageAlloc(g, g->dStageWl, g->stageWl, nchan, "stage wavelengths"));
  HWS_TRY(stageAlloc(g, g->dStageVis, g->stageVis, visFloats, "stage vis"));
  HWS_TRY(checkHip(g, hipMemcpyAsync(g->dStageUvw, uvw, nrow * 3 * sizeof(double),
                                     hipMemcpyHostToDevice, g->stream), "uvw upload"));
  return checkHip(g, hipMemcpyAsync(g->dStageWl, wl, nchan * sizeof(double),
                                    hipMemcpyHostToDevice, g->stream), "wavelength upload");
}

hws_status record(hws_gridder* g, int i) {
  if (g->capturing) return HWS_OK;  // phase timings are not part of a captured graph
  return checkHip(g, hipEventRecord(g->ev[i], g->stream), "event record");
}

}  // namespace

extern "C" {

const char* hws_status_string(hws_status s) {
  switch (s) {
    case HWS_OK: return "ok";
    case HWS_ERR_INVALID_ARGUMENT: return "invalid argument";
    case HWS_ERR_HIP: return "HIP error";
    case HWS_ERR_OUT_OF_MEMORY: return "out of device memory";
    case HWS_ERR_UNSUPPORTED: return "unsupported";
    case HWS_ERR_BAD_STATE: return "call sequence violated";
  }
  return "unknown status";
}

const char* hws_last_error(const hws_gridder* g) {
  return g ? g->err.c_str() : threadError().c_str();
}

hws_status hws_get_kernel(int kernel_mode, size_t oversampling, size_t size,
                          double* kernel) {
  if (!kernel || oversampling == 0 || size == 0)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "invalid kernel request");
  std::vector<double> k;
  if (!makeKernel(kernel_mode, oversampling, size, true, k))
    return fail(nullptr, HWS_ERR_UNSUPPORTED,
                "Blackman-Harris kernel not supported by W-Stacking gridder");
  std::memcpy(kernel, k.data(), k.size() * sizeof(double));
  return HWS_OK;
}

hws_status hws_create(const hws_config* c, hws_gridder** out) {
  if (!c || !out) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (!validSize(c->width) || !validSize(c->height))
    return fail(nullptr, HWS_ERR_UNSUPPORTED,
                "width and height must be even products of 2, 3, 5, 7 up to 16384");
  if (c->kernel_size == 0 || c->kernel_size % 2 == 0 || c->kernel_size > kMaxKernelSize)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "kernel size must be odd and at most 15");
  if (c->oversampling == 0 || c->oversampling > kMaxOversampling)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "oversampling must be 1..1023");
  if (c->kernel_mode == HWS_KERNEL_BLACKMAN_HARRIS)
    return fail(nullptr, HWS_ERR_UNSUPPORTED,
                "Blackman-Harris kernel not supported by W-Stacking gridder");
  if (c->kernel_mode < 0 || c->kernel_mode > HWS_KERNEL_BLACKMAN_HARRIS)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "unknown kernel mode");
  if (c->precision < HWS_PRECISION_PLAIN_FP32 || c->precision > HWS_PRECISION_MIXED)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "unknown precision mode");
  if (c->n_images < 1 || c->n_images > size_t(kMaxImages))
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "n_images must be 1..4");
  if (c->tile_size != 0 && (c->tile_size < 8 || c->tile_size > 64))
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "tile_size must be 0 or 8..64");
  if (c->work_chunk > (size_t(1) << 30))
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "work_chunk must be at most 2^30");
  if (!(c->pixel_size_x > 0.0) || !(c->pixel_size_y > 0.0))
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "pixel sizes must be positive");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, HWS_ERR_HIP, "no HIP device available");
  if (c->device < 0 || c->device >= ndev)
    return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "device ordinal out of range");

  hws_gridder* g = new (std::nothrow) hws_gridder();
  if (!g) return fail(nullptr, HWS_ERR_OUT_OF_MEMORY, "host allocation failed");
  g->cfg = *c;
  g->W = int(c->width);
  g->H = int(c->height);
  g->nImg = int(c->n_images);
  for (int i = 0; i < g->nImg; ++i) {
    g->isComplex[i] = (c->complex_mask >> i) & 1u;
    (g->isComplex[i] ? g->anyComplex : g->anyReal) = true;
  }
  const bool nn = c->kernel_mode == HWS_KERNEL_NEAREST_NEIGHBOUR;
  // nearest neighbour deposits on one cell: a 1-tap kernel of value 1
  g->K = nn ? 1 : int(c->kernel_size);
  g->OS = nn ? 1 : int(c->oversampling);

  auto bail = [&](hws_status s) {
    const std::string msg = g->err;
    hws_destroy(g);
    threadError() = msg;
    return s;
  };
  if (hipSetDevice(c->device) != hipSuccess) return bail(fail(g, HWS_ERR_HIP, "hipSetDevice failed"));
  hws_status st = checkHip(g, hipStreamCreateWithFlags(&g->ownStream, hipStreamNonBlocking), "stream");
  if (st != HWS_OK) return bail(st);
  g->stream = g->ownStream;
  for (auto& e : g->ev)
    if ((st = checkHip(g, hipEventCreate(&e), "event")) != HWS_OK) return bail(st);

  // kernel tables: tables[OS-1-i][x] = k1d[x*OS + i] (cpp:374-384)
  makeKernel(c->kernel_mode, c->oversampling, c->kernel_size, false, g->k1d);
  std::vector<float> tab(size_t(g->OS) * g->K);
  std::vector<__half2> pair(tab.size());
  if (nn) {
    tab[0] = 1.0f;
  } else {
    for (int i = 0; i < g->OS; ++i)
      for (int x = 0; x < g->K; ++x)
        tab[size_t(g->OS - i - 1) * g->K + x] = float(g->k1d[size_t(x) * g->OS + i]);
  }
  float tabMax = 0.0f;
  for (float v : tab) tabMax = std::max(tabMax, std::fabs(v));
  (void)std::frexp(tabMax, &g->kernelExp);
  for (size_t i = 0; i < tab.size(); ++i) {
    const __half hi = __float2half(tab[i]);
    const __half lo = __float2half(tab[i] - __half2float(hi));
    pair[i] = __halves2half2(hi, lo);
  }
  if ((st = devAlloc(g, g->dKernel, tab.size(), "kernel table")) != HWS_OK) return bail(st);
  if ((st = devAlloc(g, g->dKernelPair, pair.size(), "kernel table")) != HWS_OK) return bail(st);
  if ((st = checkHip(g, hipMemcpy(g->dKernel, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice), "kernel upload")) != HWS_OK) return bail(st);
  if ((st = checkHip(g, hipMemcpy(g->dKernelPair, pair.data(), pair.size() * sizeof(__half2), hipMemcpyHostToDevice), "kernel upload")) != HWS_OK) return bail(st);

  // kernel correction, rounded to float as the fftwf REDFT01 output is
  g->corrX = kernelCorrection(g->k1d, c->kernel_size, c->oversampling, c->width);
  g->corrY = (c->width == c->height) ? g->corrX
                                     : kernelCorrection(g->k1d, c->kernel_size, c->oversampling, c->height);
  std::vector<float> cx(g->corrX.begin(), g->corrX.end()), cy(g->corrY.begin(), g->corrY.end());
  if ((st = devAlloc(g, g->dCorrXf, cx.size(), "correction")) != HWS_OK) return bail(st);
  if ((st = devAlloc(g, g->dCorrYf, cy.size(), "correction")) != HWS_OK) return bail(st);
  if ((st = checkHip(g, hipMemcpy(g->dCorrXf, cx.data(), cx.size() * sizeof(float), hipMemcpyHostToDevice), "correction upload")) != HWS_OK) return bail(st);
  if ((st = checkHip(g, hipMemcpy(g->dCorrYf, cy.data(), cy.size() * sizeof(float), hipMemcpyHostToDevice), "correction upload")) != HWS_OK) return bail(st);

  // blocked layer layout: always at the power-of-two sizes with compile-time
  // transforms; at other sizes for the binned forms only (the plain form's
  // per-tap global atomics touch fewer cache lines in row-major order)
  g->blocked = blockedLayoutSupported(g->W, g->H) &&
               ((fastBlockedSize(g->W) && fastBlockedSize(g->H)) ||
                c->precision != HWS_PRECISION_PLAIN_FP32);
  {
    const char* e = std::getenv("HWS_NO_MFMA");
    g->noMfma = e && e[0] == '1';
    const char* f = std::getenv("HWS_NO_FUSED_PROJECT");
    g->noFusedProject = f && f[0] == '1';
    const char* h = std::getenv("HWS_NO_F16_MFMA");
    g->noF16Mfma = h && h[0] == '1';
    const char* tl = std::getenv("HWS_FFT_LINES");
    if (tl && (std::atoi(tl) == 1 || std::atoi(tl) == 2 || std::atoi(tl) == 4))
      g->fftLines = std::atoi(tl);
    const char* d = std::getenv("HWS_DEGRID_BLOCKS_PER_CU");
    if (d && std::atoi(d) > 0 && std::atoi(d) <= 16) g->degridBlocksPerCu = std::atoi(d);
  }
  // uv tile edge of the binned forms: divides W and H, a multiple of 4, at
  // least K - 1 (a footprint then spans at most two tiles per axis) and with
  // T + K - 1 <= 64 (6-bit record offsets); 16 (the matrix-core grid
  // kernel's tile) unless the caller names one
  {
    const int K = g->K;
    const int want = c->tile_size ? int(c->tile_size) : 16;
    int best = 0;
    for (int t = 4; t + K - 1 <= 64 && t <= 64; t += 4)
      if (g->W % t == 0 && g->H % t == 0 && t >= K - 1 && std::abs(t - want) <= std::abs(best - want))
        best = t;
    if (c->tile_size && best != want) {
      threadError() = "tile_size must divide width and height, be a multiple of 4 and >= K - 1";
      hws_destroy(g);
      return HWS_ERR_INVALID_ARGUMENT;
    }
    if (best == 0 && c->precision != HWS_PRECISION_PLAIN_FP32) {
      threadError() = "no uv tile size divides both image dimensions";
      hws_destroy(g);
      return HWS_ERR_UNSUPPORTED;
    }
    g->tileSize = best ? best : 32;
  }
  if ((st = makeFftPlan(g, g->W, g->rowPlan)) != HWS_OK) return bail(st);
  if ((st = makeFftPlan(g, g->H, g->colPlan)) != HWS_OK) return bail(st);
  if (g->blocked && !(fastBlockedSize(g->W) && fastBlockedSize(g->H)) && !blockedGenericFits(g))
    g->blocked = false;
  if ((st = buildF16Table(g)) != HWS_OK) return bail(st);
  *out = g;
  return HWS_OK;
}

void hws_destroy(hws_gridder* g) {
  if (!g) return;
  g->generation->alive = false;  // graphs of this context refuse to launch
  if (g->cfg.device >= 0) (void)hipSetDevice(g->cfg.device);
  if (g->ownStream) (void)hipStreamSynchronize(g->ownStream);
  if (g->userStream) (void)hipStreamSynchronize(g->stream);
  freeDev(g->dKernel);
  freeDev(g->dKernelPair);
  freeDev(g->dF16Table);
  freeDev(g->dCorrXf);
  freeDev(g->dCorrYf);
  freeDev(g->dLayers);
  freeDev(g->dSum);
  freeDev(g->dComp);
  freeDev(g->dISum);
  freeDev(g->dIComp);
  freeDev(g->dModelRe);
  freeDev(g->dModelIm);
  freeDev(g->dNTable);
  freeDev(g->dStageUvw);
  freeDev(g->dStageWl);
  freeDev(g->dStageVis);
  freeDev(g->dStageOut);
  freeStaging(g);
  freeFftPlan(g->rowPlan);
  freeFftPlan(g->colPlan);
  Plan& pl = g->plan;
  freeDev(pl.scanTmp);
  freeDev(pl.scanScratch);
  freeDev(pl.counters);
  freeDev(pl.queue);
  freeDev(pl.rot);
  freeDev(pl.cw);
  freeDev(g->dWeightSums);
  freeDev(g->dWeightPartials);
  freeDev(pl.rowKeys);
  freeDev(pl.rowKeysAlt);
  freeDev(pl.rowOrder);
  freeDev(pl.rowOrderAlt);
  for (auto& kp : pl.kinds) {
    freeDev(kp.recSorted);
    freeDev(kp.counts);
    freeDev(kp.binStart);
    freeDev(kp.extraStart);
    freeDev(kp.itemStart);
    freeDev(kp.itemBin);
    freeDev(kp.occ);
    freeDev(kp.need);
    freeDev(kp.rowLive);
    freeDev(kp.colNeed);
    freeDev(kp.partials);
    freeDev(kp.skipped);
    freeDev(kp.keys);
    freeDev(kp.keysAlt);
    freeDev(kp.recAlt);
  }
  if (pl.sortTemp) (void)hipFree(pl.sortTemp);
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  if (g->ownStream) (void)hipStreamDestroy(g->ownStream);
  delete g;
}

hws_status hws_set_stream(hws_gridder* g, void* hip_stream) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  // NULL is HIP's null stream, as everywhere in HIP
  g->stream = static_cast<hipStream_t>(hip_stream);
  g->userStream = true;
  return HWS_OK;
}

hws_status hws_synchronize(hws_gridder* g) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  return checkHip(g, hipStreamSynchronize(g->stream), "synchronize");
}

hws_status hws_prepare_w_layers(hws_gridder* g, size_t n_w_layers,
                                double max_mem, double min_w, double max_w,
                                size_t* n_passes) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (n_w_layers == 0 || n_w_layers > 65536)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "number of w-layers must be 1..65536");
  if (!std::isfinite(min_w) || !std::isfinite(max_w))
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "w limits must be finite");
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "synchronize"));
  g->minW = min_w;
  g->maxW = max_w;
  // gridding/wstackinggridder.cpp:70-75
  if (g->minW == g->maxW) g->maxW += 1.0;
  g->nW = int(n_w_layers);
  // the plan's bins follow the layer map, and the layer buffers may move:
  // plan again, capture again
  g->plan.valid = false;
  ++g->generation->value;
  const size_t HW = size_t(g->W) * g->H;
  const size_t perLayer = layersBytes(g, 1);
  double budget = max_mem;
  if (!(budget > 0.0)) {
    size_t freeMem = 0, totalMem = 0;
    HWS_TRY(checkHip(g, hipMemGetInfo(&freeMem, &totalMem), "hipMemGetInfo"));
    // keep room for the image accumulators and a binning plan
    const double reserve = double(8 * HW * g->nImg * sizeof(float)) + 0.15 * double(freeMem);
    budget = 0.75 * double(freeMem) - reserve;
  }
  size_t perPass = size_t(std::max(1.0, std::floor(budget / double(perLayer))));
  perPass = std::min<size_t>(perPass, n_w_layers);
  g->nPasses = int((n_w_layers + perPass - 1) / perPass);
  int maxL = 0;
  for (int p = 0; p < g->nPasses; ++p) maxL = std::max(maxL, layersInPass(g, p));
  g->maxLayersPerPass = maxL;
  HWS_TRY(devAlloc(g, g->dLayers, layersBytes(g, maxL) / sizeof(float2), "w-layers"));
  g->layersZero = false;
  HWS_TRY(devAlloc(g, g->dSum, g->nImg * HW, "image accumulator"));
  HWS_TRY(devAlloc(g, g->dComp, g->nImg * HW, "image accumulator"));
  if (g->anyComplex) {
    HWS_TRY(devAlloc(g, g->dISum, g->nImg * HW, "image accumulator"));
    HWS_TRY(devAlloc(g, g->dIComp, g->nImg * HW, "image accumulator"));
  }
  HWS_TRY(zeroAccumulators(g));
  g->curPass = -1;
  g->inversionActive = g->predictionActive = false;
  if (n_passes) *n_passes = size_t(g->nPasses);
  return HWS_OK;
}

size_t hws_w_to_layer(const hws_gridder* g, double w, int is_complex) {
  if (!g || g->nW == 0) return std::numeric_limits<size_t>::max();
  if (g->nW == 1) return 0;
  double r;
  if (is_complex)
    r = std::round((w + g->maxW) * (g->nW - 1) / (g->maxW + g->maxW));
  else
    r = std::round((std::fabs(w) - g->minW) * (g->nW - 1) / (g->maxW - g->minW));
  if (!(r >= 0.0)) return std::numeric_limits<size_t>::max();
  return size_t(r);
}

double hws_layer_to_w(const hws_gridder* g, size_t layer, int is_complex) {
  if (!g) return 0.0;
  return layerToW(layerMap(g, is_complex ? 1 : 0), int(layer));
}

hws_status hws_pass_layer_range(const hws_gridder* g, size_t pass,
                                size_t* start, size_t* end) {
  if (!g || !start || !end) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null argument");
  if (g->nPasses == 0 || pass >= size_t(g->nPasses))
    return fail(const_cast<hws_gridder*>(g), HWS_ERR_INVALID_ARGUMENT, "pass out of range");
  *start = size_t(passLayerStart(g, int(pass)));
  *end = size_t(passLayerStart(g, int(pass) + 1));
  return HWS_OK;
}

hws_status hws_start_inversion_pass(hws_gridder* g, size_t pass) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!prepared(g)) return fail(g, HWS_ERR_BAD_STATE, "hws_prepare_w_layers has not been called");
  if (pass >= size_t(g->nPasses)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "pass out of range");
  (void)hipSetDevice(g->cfg.device);
  if (pass == 0 || !g->inversionActive) HWS_TRY(zeroAccumulators(g));
  HWS_TRY(zeroLayers(g));
  g->curPass = int(pass);
  g->inversionActive = true;
  g->predictionActive = false;
  return HWS_OK;
}

static hws_status addRows(hws_gridder* g, size_t nrow, size_t nchan,
                          const double* uvw, const double* wl, const float* vis) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->inversionActive || g->curPass < 0)
    return fail(g, HWS_ERR_BAD_STATE, "no inversion pass started");
  if (nrow == 0 || nchan == 0) return HWS_OK;
  if (!uvw || !wl || !vis) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  (void)hipSetDevice(g->cfg.device);
  const size_t nvis = nrow * nchan * size_t(g->nImg) * 2;
  HWS_TRY(stageRows(g, nrow, nchan, uvw, wl, nvis));
  HWS_TRY(checkHip(g, hipMemcpyAsync(g->dStageVis, vis, nvis * sizeof(float),
                                     hipMemcpyHostToDevice, g->stream), "vis upload"));
  g->layersZero = false;
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
    HWS_TRY(gridPlain(g, nrow, nchan, g->dStageUvw, g->dStageWl, g->dStageVis, g->curPass));
  } else {
    HWS_TRY(buildPlan(g, nrow, nchan, g->dStageUvw, g->dStageWl));
    HWS_TRY(gridBinned(g, g->dStageVis, g->curPass, true));
  }
  return checkHip(g, hipStreamSynchronize(g->stream), "add data");
}

hws_status hws_add_data(hws_gridder* g, size_t nrow, size_t nchan,
                        const double* uvw_m, const double* wavelengths,
                        const float* vis) {
  return addRows(g, nrow, nchan, uvw_m, wavelengths, vis);
}

hws_status hws_add_samples(hws_gridder* g, size_t n, const double* uvw_lambda,
                           const float* vis) {
  const double one = 1.0;
  return addRows(g, n, 1, uvw_lambda, &one, vis);
}

hws_status hws_finish_inversion_pass(hws_gridder* g) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->inversionActive || g->curPass < 0)
    return fail(g, HWS_ERR_BAD_STATE, "no inversion pass started");
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(fftPassLayers(g, g->curPass, +1));
  HWS_TRY(projectPass(g, g->curPass));
  g->layersZero = false;
  return checkHip(g, hipStreamSynchronize(g->stream), "finish inversion pass");
}

hws_status hws_finalize_image(hws_gridder* g, double mult, float* real,
                              float* imag) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->inversionActive) return fail(g, HWS_ERR_BAD_STATE, "no inversion to finalise");
  if (!real) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  (void)hipSetDevice(g->cfg.device);
  const size_t n = size_t(g->nImg) * g->W * g->H;
  HWS_TRY(stageAlloc(g, g->dStageOut, g->stageOut, 2 * n, "stage image"));
  float* dRe = g->dStageOut;
  float* dIm = g->dStageOut + n;
  HWS_TRY(finalizeImages(g, mult, dRe, (imag && g->anyComplex) ? dIm : nullptr));
  HWS_TRY(checkHip(g, hipMemcpyAsync(real, dRe, n * sizeof(float), hipMemcpyDeviceToHost, g->stream), "image download"));
  if (imag && g->anyComplex)
    HWS_TRY(checkHip(g, hipMemcpyAsync(imag, dIm, n * sizeof(float), hipMemcpyDeviceToHost, g->stream), "image download"));
  g->inversionActive = false;
  return checkHip(g, hipStreamSynchronize(g->stream), "finalize");
}

static hws_status ensureModel(hws_gridder* g) {
  const size_t n = size_t(g->nImg) * g->W * g->H;
  if (!g->dModelRe) {
    HWS_TRY(devAlloc(g, g->dModelRe, n, "model image"));
    HWS_TRY(launchFill(g, g->dModelRe, 0u, n));
  }
  if (g->anyComplex && !g->dModelIm) {
    HWS_TRY(devAlloc(g, g->dModelIm, n, "model image"));
    HWS_TRY(launchFill(g, g->dModelIm, 0u, n));
  }
  return HWS_OK;
}

hws_status hws_initialize_prediction(hws_gridder* g, const float* real,
                                     const float* imag) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!prepared(g)) return fail(g, HWS_ERR_BAD_STATE, "hws_prepare_w_layers has not been called");
  if (!real) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  if (g->anyComplex && !imag)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "Missing imaginary in complex prediction");
  (void)hipSetDevice(g->cfg.device);
  const size_t n = size_t(g->nImg) * g->W * g->H;
  HWS_TRY(ensureModel(g));
  HWS_TRY(stageAlloc(g, g->dStageOut, g->stageOut, 2 * n, "stage image"));
  HWS_TRY(checkHip(g, hipMemcpyAsync(g->dStageOut, real, n * sizeof(float), hipMemcpyHostToDevice, g->stream), "model upload"));
  if (g->anyComplex)
    HWS_TRY(checkHip(g, hipMemcpyAsync(g->dStageOut + n, imag, n * sizeof(float), hipMemcpyHostToDevice, g->stream), "model upload"));
  HWS_TRY(initPrediction(g, g->dStageOut, g->anyComplex ? g->dStageOut + n : nullptr));
  g->predictionActive = true;
  g->inversionActive = false;
  g->curPass = -1;
  return checkHip(g, hipStreamSynchronize(g->stream), "initialize prediction");
}

hws_status hws_start_prediction_pass(hws_gridder* g, size_t pass) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->predictionActive) return fail(g, HWS_ERR_BAD_STATE, "prediction not initialised");
  if (pass >= size_t(g->nPasses)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "pass out of range");
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(imageToLayers(g, int(pass)));
  HWS_TRY(fftPassLayers(g, int(pass), -1));
  g->layersZero = false;
  g->curPass = int(pass);
  return checkHip(g, hipStreamSynchronize(g->stream), "start prediction pass");
}

static hws_status sampleRows(hws_gridder* g, size_t nrow, size_t nchan,
                             const double* uvw, const double* wl, float* out) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->predictionActive || g->curPass < 0)
    return fail(g, HWS_ERR_BAD_STATE, "no prediction pass started");
  if (nrow == 0 || nchan == 0) return HWS_OK;
  if (!uvw || !wl || !out) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  (void)hipSetDevice(g->cfg.device);
  const size_t nvis = nrow * nchan * size_t(g->nImg) * 2;
  HWS_TRY(stageRows(g, nrow, nchan, uvw, wl, nvis));
  HWS_TRY(launchFill(g, g->dStageVis, 0x7FC00000u, nvis));  // quiet NaN
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
    HWS_TRY(degridPlain(g, nrow, nchan, g->dStageUvw, g->dStageWl, g->dStageVis, g->curPass));
  } else {
    HWS_TRY(buildPlan(g, nrow, nchan, g->dStageUvw, g->dStageWl));
    HWS_TRY(degridBinned(g, g->dStageVis, g->curPass));
  }
  HWS_TRY(checkHip(g, hipMemcpyAsync(out, g->dStageVis, nvis * sizeof(float),
                                     hipMemcpyDeviceToHost, g->stream), "vis download"));
  return checkHip(g, hipStreamSynchronize(g->stream), "sample data");
}

hws_status hws_sample_data(hws_gridder* g, size_t nrow, size_t nchan,
                           const double* uvw_m, const double* wavelengths,
                           float* vis_out) {
  return sampleRows(g, nrow, nchan, uvw_m, wavelengths, vis_out);
}

hws_status hws_sample_samples(hws_gridder* g, size_t n,
                              const double* uvw_lambda, float* vis_out) {
  const double one = 1.0;
  return sampleRows(g, n, 1, uvw_lambda, &one, vis_out);
}

hws_status hws_get_gridded_uv_layer(hws_gridder* g, size_t image, size_t layer,
                                    float* out) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!out || image >= size_t(g->nImg)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "bad image or output");
  if (g->curPass < 0) return fail(g, HWS_ERR_BAD_STATE, "no pass started");
  if (layer >= size_t(layersInPass(g, g->curPass))) return fail(g, HWS_ERR_INVALID_ARGUMENT, "layer out of range");
  (void)hipSetDevice(g->cfg.device);
  const size_t HW = size_t(g->W) * g->H;
  const float2* src = g->dLayers + (image * g->maxLayersPerPass + layer) * HW;
  if (!g->blocked) {
    HWS_TRY(checkHip(g, hipMemcpyAsync(out, src, HW * sizeof(float2), hipMemcpyDeviceToHost, g->stream), "layer download"));
    return checkHip(g, hipStreamSynchronize(g->stream), "layer download");
  }
  std::vector<float2> tmp(HW);
  HWS_TRY(checkHip(g, hipMemcpyAsync(tmp.data(), src, HW * sizeof(float2), hipMemcpyDeviceToHost, g->stream), "layer download"));
  HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "layer download"));
  for (int y = 0; y < g->H; ++y)
    for (int x = 0; x < g->W; ++x) {
      const float2 v = tmp[cellIndex(x, y, g->W, true)];
      out[2 * (size_t(y) * g->W + x)] = v.x;
      out[2 * (size_t(y) * g->W + x) + 1] = v.y;
    }
  return HWS_OK;
}

hws_status hws_plan_device(hws_gridder* g, size_t nrow, size_t nchan,
                           const double* d_uvw_m, const double* d_wavelengths) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!prepared(g)) return fail(g, HWS_ERR_BAD_STATE, "hws_prepare_w_layers has not been called");
  if ((nrow && !d_uvw_m) || (nchan && !d_wavelengths)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(record(g, 0));
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
    Plan& pl = g->plan;
    ++g->generation->value;
    pl.nrow = nrow;
    pl.nchan = nchan;
    pl.uvw = d_uvw_m;
    pl.wavelengths = d_wavelengths;
    pl.valid = true;
  } else {
    HWS_TRY(buildPlan(g, nrow, nchan, d_uvw_m, d_wavelengths));
  }
  g->plan.device = true;
  HWS_TRY(planWeights(g));
  HWS_TRY(record(g, 1));
  g->stats = hws_stats{};
  g->lastOp = 0;
  g->stats.samples_in = uint64_t(nrow) * nchan;
  return HWS_OK;
}

hws_status hws_invert_planned_device(hws_gridder* g, const float* d_vis,
                                     double mult, float* d_real, float* d_imag) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->plan.valid) return fail(g, HWS_ERR_BAD_STATE, "hws_plan_device has not been called");
  if (!g->plan.device) return fail(g, HWS_ERR_BAD_STATE, kReplacedPlan);
  if (!d_real || (g->plan.nrow && !d_vis)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  (void)hipSetDevice(g->cfg.device);
  Plan& pl = g->plan;
  // weighting attached after the plan: its per-sample weights and sums first
  if (g->weights && !pl.weighted) {
    if (g->capturing)
      return fail(g, HWS_ERR_BAD_STATE, "plan again after attaching weights before a capture");
    HWS_TRY(planWeights(g));
  }
  HWS_TRY(record(g, 2));
  // the batched forms' first projection writes the accumulators outright
  const bool plainForm = g->cfg.precision == HWS_PRECISION_PLAIN_FP32;
  if (plainForm) HWS_TRY(zeroAccumulators(g));
  float ms[3] = {0, 0, 0};
  for (int pass = 0; pass < g->nPasses; ++pass) {
    if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32) {
      HWS_TRY(zeroLayers(g));
      g->layersZero = false;
      HWS_TRY(gridPlain(g, pl.nrow, pl.nchan, pl.uvw, pl.wavelengths, d_vis, pass));
    } else {
      // the non-empty tiles are written (and the rest cleared unless the
      // transform skips them)
      g->layersZero = false;
      HWS_TRY(gridBinned(g, d_vis, pass, false, sparseInvert(g)));
    }
    HWS_TRY(record(g, 3));
    const bool sparse = !plainForm && sparseInvert(g);
    // the columns pass and the projection in one kernel where it applies
    const bool fused = sparse && colsProjectSupported(g, layersInPass(g, pass));
    HWS_TRY(fftPassLayers(g, pass, +1, sparse, false, fused));
    HWS_TRY(record(g, 4));
    if (fused)
      HWS_TRY(projectFusedPass(g, pass, pass == 0));
    else
      HWS_TRY(projectPass(g, pass, !plainForm && pass == 0));
    HWS_TRY(record(g, 5));
    if (g->nPasses > 1) {
      float a = 0, b = 0, c = 0;
      HWS_TRY(checkHip(g, hipEventSynchronize(g->ev[5]), "event sync"));
      (void)hipEventElapsedTime(&a, g->ev[2], g->ev[3]);
      (void)hipEventElapsedTime(&b, g->ev[3], g->ev[4]);
      (void)hipEventElapsedTime(&c, g->ev[4], g->ev[5]);
      ms[0] += a;
      ms[1] += b;
      ms[2] += c;
      HWS_TRY(record(g, 2));
    }
  }
  HWS_TRY(finalizeImages(g, mult, d_real, d_imag));
  HWS_TRY(record(g, 6));
  g->stats.ms_grid = ms[0];
  g->stats.ms_fft = ms[1];
  g->stats.ms_project = ms[2];
  g->lastOp = 1;
  g->curPass = -1;
  g->inversionActive = false;
  return HWS_OK;
}

hws_status hws_predict_planned_device(hws_gridder* g, const float* d_real,
                                      const float* d_imag, float* d_vis_out) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->plan.valid) return fail(g, HWS_ERR_BAD_STATE, "hws_plan_device has not been called");
  if (!g->plan.device) return fail(g, HWS_ERR_BAD_STATE, kReplacedPlan);
  if (!d_real || (g->plan.nrow && !d_vis_out)) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  if (g->anyComplex && !d_imag)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "Missing imaginary in complex prediction");
  (void)hipSetDevice(g->cfg.device);
  Plan& pl = g->plan;
  HWS_TRY(record(g, 2));
  HWS_TRY(ensureModel(g));
  HWS_TRY(initPrediction(g, d_real, d_imag));
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32)
    HWS_TRY(launchFill(g, d_vis_out, 0x7FC00000u, pl.nrow * pl.nchan * size_t(g->nImg) * 2));
  else
    HWS_TRY(markSkipped(g, d_vis_out));  // the degridding writes every other value
  float ms[3] = {0, 0, 0};
  for (int pass = 0; pass < g->nPasses; ++pass) {
    // blocked binned form: the forward row pass computes the layers from the
    // model and the column pass writes only the tiles degridding reads
    const bool sparse = g->cfg.precision != HWS_PRECISION_PLAIN_FP32 && sparseInvert(g);
    if (!sparse) HWS_TRY(imageToLayers(g, pass));
    HWS_TRY(record(g, 3));
    HWS_TRY(fftPassLayers(g, pass, -1, sparse, sparse));
    HWS_TRY(record(g, 4));
    if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32)
      HWS_TRY(degridPlain(g, pl.nrow, pl.nchan, pl.uvw, pl.wavelengths, d_vis_out, pass));
    else
      HWS_TRY(degridBinned(g, d_vis_out, pass));
    HWS_TRY(record(g, 5));
    if (g->nPasses > 1) {
      float a = 0, b = 0, c = 0;
      HWS_TRY(checkHip(g, hipEventSynchronize(g->ev[5]), "event sync"));
      (void)hipEventElapsedTime(&a, g->ev[2], g->ev[3]);
      (void)hipEventElapsedTime(&b, g->ev[3], g->ev[4]);
      (void)hipEventElapsedTime(&c, g->ev[4], g->ev[5]);
      ms[0] += c;  // degrid
      ms[1] += b;  // fft
      ms[2] += a;  // image -> layers
      HWS_TRY(record(g, 2));
    }
  }
  HWS_TRY(record(g, 6));
  g->layersZero = false;
  g->stats.ms_grid = ms[0];
  g->stats.ms_fft = ms[1];
  g->stats.ms_project = ms[2];
  g->lastOp = 2;
  g->curPass = -1;
  return HWS_OK;
}

hws_status hws_invert_device(hws_gridder* g, size_t nrow, size_t nchan,
                             const double* d_uvw_m, const double* d_wavelengths,
                             const float* d_vis, double mult, float* d_real,
                             float* d_imag) {
  HWS_TRY(hws_plan_device(g, nrow, nchan, d_uvw_m, d_wavelengths));
  return hws_invert_planned_device(g, d_vis, mult, d_real, d_imag);
}

hws_status hws_predict_device(hws_gridder* g, size_t nrow, size_t nchan,
                              const double* d_uvw_m, const double* d_wavelengths,
                              const float* d_real, const float* d_imag,
                              float* d_vis_out) {
  HWS_TRY(hws_plan_device(g, nrow, nchan, d_uvw_m, d_wavelengths));
  return hws_predict_planned_device(g, d_real, d_imag, d_vis_out);
}

hws_status hws_get_stats(const hws_gridder* cg, hws_stats* out) {
  hws_gridder* g = const_cast<hws_gridder*>(cg);
  if (!g || !out) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null argument");
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "synchronize"));
  hws_stats s = g->stats;
  float bin = 0, total = 0;
  (void)hipEventElapsedTime(&bin, g->ev[0], g->ev[1]);
  s.ms_bin = bin;
  if (g->nPasses == 1 && g->lastOp != 0) {
    float a = 0, b = 0, c = 0;
    (void)hipEventElapsedTime(&a, g->ev[2], g->ev[3]);
    (void)hipEventElapsedTime(&b, g->ev[3], g->ev[4]);
    (void)hipEventElapsedTime(&c, g->ev[4], g->ev[5]);
    // invert: grid, fft, project; predict: image->layers, fft, degrid
    s.ms_grid = g->lastOp == 1 ? a : c;
    s.ms_fft = b;
    s.ms_project = g->lastOp == 1 ? c : a;
  }
  if (g->lastOp != 0) {
    (void)hipEventElapsedTime(&total, g->ev[2], g->ev[6]);
    s.ms_total = total;
  }
  if (g->plan.counters) {
    unsigned long long c[kPlanCounters] = {};
    HWS_TRY(checkHip(g, hipMemcpyAsync(c, g->plan.counters, sizeof(c), hipMemcpyDeviceToHost,
                                       g->stream), "counters"));
    HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "counters"));
    // summed over the image kinds present (each kind places every sample)
    s.samples_gridded = c[0] + c[3];
    s.samples_off_grid = c[1] + c[4];
    s.samples_off_layer = c[2] + c[5];
    s.bins_used = c[6] + c[7];
    s.tiles_total = g->plan.nbins;
    for (int k = 0; k < 2; ++k) {
      s.tiles_occupied[k] = c[8 + k];
      s.tiles_needed[k] = c[10 + k];
      s.tile_rows_live[k] = c[14 + k];
      s.tile_cols_needed[k] = c[16 + k];
    }
    s.tile_rows_total = g->plan.tilesX ? g->plan.nbins / size_t(g->plan.tilesX) : 0;
    s.tile_cols_total = g->plan.tilesY ? g->plan.nbins / size_t(g->plan.tilesY) : 0;
    for (const auto& kp : g->plan.kinds)
      if (kp.active) s.chunks += kp.nWorkItems;
  }
  if (g->weights && g->weightSumsValid) {
    double w[kMaxImages * 4] = {};
    HWS_TRY(checkHip(g, hipMemcpyAsync(w, g->dWeightSums, sizeof(w), hipMemcpyDeviceToHost,
                                       g->stream), "weight sums"));
    HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "weight sums"));
    for (int i = 0; i < g->nImg; ++i) {
      s.total_weight[i] = w[i * 4];
      s.visibility_weight_sum[i] = w[i * 4 + 1];
      s.max_gridded_weight[i] = w[i * 4 + 2];
      s.gridded_visibility_count[i] = uint64_t(w[i * 4 + 3]);
    }
  }
  *out = s;
  return HWS_OK;
}

hws_status hws_set_weighting(hws_gridder* g, const hws_weights* weights,
                             const float* d_vis_weights, int visibility_weighting_mode) {
  if (!g) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (visibility_weighting_mode < 0 || visibility_weighting_mode > 2)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "unknown visibility weighting mode");
  if (weights && !weights->finished)
    return fail(g, HWS_ERR_BAD_STATE,
                "imaging weights are not finished (call hws_weights_finish first)");
  if (weights && weights->device != g->cfg.device)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "imaging weights live on another device");
  if (!weights && d_vis_weights)
    return fail(g, HWS_ERR_INVALID_ARGUMENT, "visibility weights need imaging weights");
  g->weights = weights;
  g->dVisWeights = weights ? d_vis_weights : nullptr;
  g->visWeightMode = weights ? visibility_weighting_mode : 0;
  ++g->generation->value;  // graphs captured with the old weighting are stale
  // the imaging weights of the samples and the weight sums are computed by
  // the next hws_plan_device, or by the next invert of the current plan
  g->plan.weighted = false;
  g->weightSumsValid = false;
  return HWS_OK;
}

// ---- graph capture of the device-resident forms
struct hws_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  std::vector<hipEvent_t> events;  // fork/join events of a batch capture
  // each captured context's plan generation at the capture
  std::vector<std::pair<std::shared_ptr<PlanGeneration>, uint64_t>> generations;
};

}  // extern "C"

namespace {

// The per-sample imaging weights and weight sums of an attached weighting,
// computed (and their buffers allocated) before a capture starts.
hws_status weightsBeforeCapture(hws_gridder* g) {
  if (!g->weights || g->plan.weighted) return HWS_OK;
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(planWeights(g));
  return checkHip(g, hipStreamSynchronize(g->stream), "weights before capture");
}

hws_status graphChecks(hws_gridder* g, hws_graph** out) {
  if (!g || !out) return fail(g, HWS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (!g->plan.valid) return fail(g, HWS_ERR_BAD_STATE, "hws_plan_device has not been called");
  if (g->nPasses != 1)
    return fail(g, HWS_ERR_UNSUPPORTED, "graph capture needs the layers in a single pass");
  if (g->cfg.precision == HWS_PRECISION_PLAIN_FP32)
    return fail(g, HWS_ERR_UNSUPPORTED, "graph capture is for the batched precisions");
  if (g->capturing) return fail(g, HWS_ERR_BAD_STATE, "a capture is already in progress");
  if (!g->plan.device) return fail(g, HWS_ERR_BAD_STATE, kReplacedPlan);
  return HWS_OK;
}

// Captures op(i) for the n contexts into one graph. Context 0 runs on its
// own stream, where the capture starts (the null stream cannot be captured);
// every other context runs on its own stream too, forked from the start of
// the capture and joined at its end, so the graph holds one independent
// branch per context and the runtime may overlap them. Each context is
// switched to its stream while it is captured.
template <class Op>
hws_status captureBatch(hws_gridder* const* gs, size_t n, Op op, hws_graph** out) {
  hws_gridder* g = gs[0];
  (void)hipSetDevice(g->cfg.device);
  hipStream_t cap = g->ownStream;
  std::vector<hipStream_t> saved(n);
  std::vector<hipEvent_t> events;
  auto dropEvents = [&] {
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  };
  for (size_t i = 0; i < n; ++i) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      dropEvents();
      return fail(g, HWS_ERR_HIP, "event creation failed");
    }
    events.push_back(ev);  // events[0] forks, events[i > 0] joins branch i
  }
  if (checkHip(g, hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal),
               "begin capture") != HWS_OK) {
    dropEvents();
    return HWS_ERR_HIP;
  }
  hipError_t ce = hipEventRecord(events[0], cap);
  for (size_t i = 0; i < n; ++i) {
    saved[i] = gs[i]->stream;
    gs[i]->stream = i == 0 ? cap : gs[i]->ownStream;
    gs[i]->capturing = true;
    if (i > 0 && ce == hipSuccess) ce = hipStreamWaitEvent(gs[i]->ownStream, events[0], 0);
  }
  hws_status st = ce == hipSuccess ? HWS_OK : fail(g, HWS_ERR_HIP, "capture fork failed");
  for (size_t i = 0; i < n && st == HWS_OK; ++i) {
    st = op(i);
    if (st != HWS_OK && gs[i] != g) g->err = gs[i]->err;
  }
  for (size_t i = 1; i < n; ++i) {
    // join every forked branch back, also after a failure, so that the
    // capture can end
    hipError_t je = hipEventRecord(events[i], gs[i]->ownStream);
    if (je == hipSuccess) je = hipStreamWaitEvent(cap, events[i], 0);
    if (je != hipSuccess && st == HWS_OK) st = fail(g, HWS_ERR_HIP, "capture join failed");
  }
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cap, &graph);
  for (size_t i = 0; i < n; ++i) {
    gs[i]->capturing = false;
    gs[i]->stream = saved[i];
  }
  if (st != HWS_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    dropEvents();
    return st;
  }
  if (checkHip(g, e, "end capture") != HWS_OK) {
    dropEvents();
    return HWS_ERR_HIP;
  }
  hipGraphExec_t exec = nullptr;
  hws_status ist = checkHip(g, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0),
                            "graph instantiate");
  if (ist != HWS_OK) {
    (void)hipGraphDestroy(graph);
    dropEvents();
    return ist;
  }
  hws_graph* gr = new (std::nothrow) hws_graph;
  if (!gr) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    dropEvents();
    return fail(g, HWS_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  gr->graph = graph;
  gr->exec = exec;
  gr->device = g->cfg.device;
  gr->events = std::move(events);
  for (size_t i = 0; i < n; ++i)
    gr->generations.emplace_back(gs[i]->generation, gs[i]->generation->value.load());
  *out = gr;
  return HWS_OK;
}

// Checks the contexts of a batch capture (each as for a single capture, all on
// one device, none twice) and, for predict, allocates their model buffers.
// A failure is reported on the first context as well as on the failing one.
hws_status batchChecks(hws_gridder* const* gs, size_t n, bool predict) {
  for (size_t i = 0; i < n; ++i) {
    if (!gs[i]) return fail(gs[0], HWS_ERR_INVALID_ARGUMENT, "null context in the batch");
    hws_graph* dummy = nullptr;
    hws_status st = graphChecks(gs[i], &dummy);
    if (st == HWS_OK && gs[i]->cfg.device != gs[0]->cfg.device)
      st = fail(gs[i], HWS_ERR_INVALID_ARGUMENT, "contexts of a batch share one device");
    for (size_t j = 0; j < i && st == HWS_OK; ++j)
      if (gs[j] == gs[i])
        st = fail(gs[i], HWS_ERR_INVALID_ARGUMENT, "a context appears twice in the batch");
    if (st == HWS_OK && predict) {
      (void)hipSetDevice(gs[i]->cfg.device);
      st = ensureModel(gs[i]);  // allocations cannot be captured
      if (st == HWS_OK)
        st = checkHip(gs[i], hipStreamSynchronize(gs[i]->stream), "model allocation");
    }
    if (st != HWS_OK) {
      if (gs[i] != gs[0]) gs[0]->err = gs[i]->err;
      return st;
    }
  }
  return HWS_OK;
}

}  // namespace

extern "C" {

hws_status hws_graph_capture_invert(hws_gridder* g, const float* d_vis, double mult,
                                    float* d_real, float* d_imag, hws_graph** out) {
  HWS_TRY(graphChecks(g, out));
  HWS_TRY(weightsBeforeCapture(g));  // allocations cannot be captured
  return captureBatch(
      &g, 1, [&](size_t) { return hws_invert_planned_device(g, d_vis, mult, d_real, d_imag); },
      out);
}

hws_status hws_graph_capture_predict(hws_gridder* g, const float* d_real, const float* d_imag,
                                     float* d_vis_out, hws_graph** out) {
  HWS_TRY(graphChecks(g, out));
  (void)hipSetDevice(g->cfg.device);
  HWS_TRY(ensureModel(g));  // allocations cannot be captured
  HWS_TRY(checkHip(g, hipStreamSynchronize(g->stream), "model allocation"));
  return captureBatch(
      &g, 1, [&](size_t) { return hws_predict_planned_device(g, d_real, d_imag, d_vis_out); },
      out);
}

hws_status hws_graph_capture_invert_batch(hws_gridder* const* gs, size_t n,
                                          const float* const* d_vis, const double* mult,
                                          float* const* d_real, float* const* d_imag,
                                          hws_graph** out) {
  if (!gs || n == 0 || !gs[0] || !d_vis || !mult || !d_real || !out)
    return fail(gs && n ? gs[0] : nullptr, HWS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  HWS_TRY(batchChecks(gs, n, false));
  for (size_t i = 0; i < n; ++i) HWS_TRY(weightsBeforeCapture(gs[i]));
  return captureBatch(gs, n, [&](size_t i) {
    return hws_invert_planned_device(gs[i], d_vis[i], mult[i], d_real[i],
                                     d_imag ? d_imag[i] : nullptr);
  }, out);
}

hws_status hws_graph_capture_predict_batch(hws_gridder* const* gs, size_t n,
                                           const float* const* d_real,
                                           const float* const* d_imag,
                                           float* const* d_vis_out, hws_graph** out) {
  if (!gs || n == 0 || !gs[0] || !d_real || !d_vis_out || !out)
    return fail(gs && n ? gs[0] : nullptr, HWS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  HWS_TRY(batchChecks(gs, n, true));
  return captureBatch(gs, n, [&](size_t i) {
    return hws_predict_planned_device(gs[i], d_real[i], d_imag ? d_imag[i] : nullptr,
                                      d_vis_out[i]);
  }, out);
}

hws_status hws_graph_launch(hws_graph* gr, void* hip_stream) {
  if (!gr) return fail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null graph");
  for (const auto& gen : gr->generations) {
    if (!gen.first->alive)
      return fail(nullptr, HWS_ERR_BAD_STATE, "a context of the graph has been destroyed");
    if (gen.first->value != gen.second)
      return fail(nullptr, HWS_ERR_BAD_STATE,
                  "the plan, w-layers or weighting of a context changed after the capture: "
                  "capture again");
  }
  (void)hipSetDevice(gr->device);
  return checkHip(nullptr, hipGraphLaunch(gr->exec, static_cast<hipStream_t>(hip_stream)),
                  "graph launch");
}

void hws_graph_destroy(hws_graph* gr) {
  if (!gr) return;
  (void)hipSetDevice(gr->device);
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  for (hipEvent_t ev : gr->events) (void)hipEventDestroy(ev);
  delete gr;
}

}  // extern "C"
