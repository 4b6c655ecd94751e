// Staged host rows for the planned device path, and the image ends of a
// WSMSGridder call on the device: the small-inversion resampling and the
// trim after an invert (gridding/wsmsgridder.cpp:452-484), the untrim and
// decimation before a predict (:510-551).
//
// A host caller (the MSGridderBase plugin) reads and weights its rows in
// batches and hands each batch to hws_stage_rows. The rows are copied into
// one of two pinned host buffers and uploaded from there asynchronously, so
// the next batch is read while the previous one crosses the bus; the upload
// waits only when both pinned buffers are in flight. hws_stage_end plans the
// staged samples once, and every pass of the invert or predict then runs on
// the device.
#include "hws_internal.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace hws {
namespace {

constexpr size_t kPinnedBytes = size_t(64) << 20;

#define HWS_STRY(expr)             \
  do {                             \
    const hws_status st_ = (expr); \
    if (st_ != HWS_OK) return st_; \
  } while (0)

hws_status sfail(hws_gridder* g, hws_status s, const std::string& m) {
  return setLastError(g, s, m);
}

template <typename T>
hws_status grow(hws_gridder* g, T*& p, size_t& cap, size_t n, const char* what) {
  if (p && cap >= n) return HWS_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return sfail(g, HWS_ERR_OUT_OF_MEMORY, std::string("out of device memory: ") + what);
  cap = n;
  return HWS_OK;
}

// memcpy split over up to eight host threads (one core's copy rate is below
// the bus's), for the copies between the caller's memory and the pinned
// buffers.
void parallelCopy(void* dst, const void* src, size_t n) {
  constexpr size_t kMinShare = size_t(4) << 20;
  const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, hc), n / kMinShare);
  if (nt < 2) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t share = ((n + nt - 1) / nt + 63) & ~size_t(63);
  auto part = [&](size_t t) {
    const size_t o = t * share;
    if (o < n)
      std::memcpy(static_cast<unsigned char*>(dst) + o, static_cast<const unsigned char*>(src) + o,
                  std::min(share, n - o));
  };
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(part, t);
  part(0);
  for (auto& th : pool) th.join();
}

hws_status ensurePinned(hws_gridder* g) {
  for (int k = 0; k < 2; ++k) {
    if (!g->pinned[k]) {
      if (hipHostMalloc(reinterpret_cast<void**>(&g->pinned[k]), kPinnedBytes,
                        hipHostMallocDefault) != hipSuccess)
        return sfail(g, HWS_ERR_OUT_OF_MEMORY, "pinned staging buffer");
      HWS_STRY(checkHip(g, hipEventCreateWithFlags(&g->pinnedFree[k], hipEventDisableTiming),
                        "staging event"));
      HWS_STRY(checkHip(g, hipEventRecord(g->pinnedFree[k], g->stream), "staging event"));
    }
  }
  return HWS_OK;
}

// Copies n bytes of host memory to dst (device) through the pinned buffers.
hws_status upload(hws_gridder* g, void* dst, const void* src, size_t n) {
  HWS_STRY(ensurePinned(g));
  const unsigned char* s = static_cast<const unsigned char*>(src);
  unsigned char* d = static_cast<unsigned char*>(dst);
  while (n > 0) {
    const int k = g->pinnedNext;
    const size_t piece = std::min(n, kPinnedBytes);
    HWS_STRY(checkHip(g, hipEventSynchronize(g->pinnedFree[k]), "staging buffer wait"));
    parallelCopy(g->pinned[k], s, piece);
    HWS_STRY(checkHip(g, hipMemcpyAsync(d, g->pinned[k], piece, hipMemcpyHostToDevice,
                                        g->stream),
                      "row upload"));
    HWS_STRY(checkHip(g, hipEventRecord(g->pinnedFree[k], g->stream), "staging event"));
    g->pinnedNext ^= 1;
    s += piece;
    d += piece;
    n -= piece;
  }
  return HWS_OK;
}

// Copies n bytes of device memory (written by work on g->stream) to host
// memory through the pinned buffers: piece i + 2 crosses the bus while piece
// i is copied out. Returns when dst holds every byte.
hws_status download(hws_gridder* g, void* dst, const void* src, size_t n) {
  HWS_STRY(ensurePinned(g));
  const unsigned char* s = static_cast<const unsigned char*>(src);
  unsigned char* d = static_cast<unsigned char*>(dst);
  const size_t pieces = (n + kPinnedBytes - 1) / kPinnedBytes;
  auto issue = [&](size_t i) -> hws_status {
    const int k = int(i & 1);
    const size_t o = i * kPinnedBytes, piece = std::min(kPinnedBytes, n - o);
    HWS_STRY(checkHip(g, hipMemcpyAsync(g->pinned[k], s + o, piece, hipMemcpyDeviceToHost,
                                        g->stream),
                      "visibility download"));
    return checkHip(g, hipEventRecord(g->pinnedFree[k], g->stream), "staging event");
  };
  for (size_t i = 0; i < std::min<size_t>(2, pieces); ++i) HWS_STRY(issue(i));
  for (size_t i = 0; i < pieces; ++i) {
    const int k = int(i & 1);
    const size_t o = i * kPinnedBytes, piece = std::min(kPinnedBytes, n - o);
    HWS_STRY(checkHip(g, hipEventSynchronize(g->pinnedFree[k]), "visibility download"));
    parallelCopy(d + o, g->pinned[k], piece);
    if (i + 2 < pieces) HWS_STRY(issue(i + 2));
  }
  return HWS_OK;
}

struct DeviceImages {
  float* p = nullptr;
  ~DeviceImages() {
    if (p) (void)hipFree(p);
  }
  hws_status alloc(hws_gridder* g, size_t n) {
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess)
      return sfail(g, HWS_ERR_OUT_OF_MEMORY, "image buffers");
    return HWS_OK;
  }
};

bool validImageSize(size_t w, size_t h) {
  return w > 0 && h > 0 && w % 2 == 0 && h % 2 == 0 && w <= 65536 && h <= 65536;
}

// The inverted images of the context size in dRe/dIm (dIm may be null) to
// host memory at the trimmed size: resampled to image_w x image_h when that
// differs from the context size, then trimmed.
hws_status finishImages(hws_gridder* g, const float* dRe, const float* dIm, size_t imageW,
                        size_t imageH, size_t trimW, size_t trimH, float* real, float* imag) {
  const size_t nImg = size_t(g->nImg);
  const float* srcs[2] = {dRe, dIm};
  float* dsts[2] = {real, imag};
  for (int part = 0; part < 2; ++part) {
    if (!srcs[part] || !dsts[part]) continue;
    const float* cur = srcs[part];
    DeviceImages resized, trimmed;
    if (imageW != size_t(g->W) || imageH != size_t(g->H)) {
      HWS_STRY(resized.alloc(g, nImg * imageW * imageH));
      HWS_STRY(hws_resample_device(g->cfg.device, g->stream, cur, size_t(g->W), size_t(g->H),
                                   resized.p, imageW, imageH, nImg));
      cur = resized.p;
    }
    if (trimW != imageW || trimH != imageH) {
      HWS_STRY(trimmed.alloc(g, nImg * trimW * trimH));
      HWS_STRY(hws_trim_device(g->cfg.device, g->stream, cur, imageW, imageH, trimmed.p, trimW,
                               trimH, nImg));
      cur = trimmed.p;
    }
    HWS_STRY(checkHip(g, hipMemcpyAsync(dsts[part], cur, nImg * trimW * trimH * sizeof(float),
                                        hipMemcpyDeviceToHost, g->stream),
                      "image download"));
    HWS_STRY(checkHip(g, hipStreamSynchronize(g->stream), "image download"));
  }
  return HWS_OK;
}

// Host models at the trimmed size to device images of the context size:
// untrimmed to image_w x image_h, then resampled to the context size. The
// results are left in out (nImg x W x H floats per part; imag part only for
// complex contexts).
hws_status prepareModels(hws_gridder* g, const float* real, const float* imag, size_t trimW,
                         size_t trimH, size_t imageW, size_t imageH, DeviceImages& outRe,
                         DeviceImages& outIm) {
  const size_t nImg = size_t(g->nImg);
  const float* srcs[2] = {real, g->anyComplex ? imag : nullptr};
  DeviceImages* outs[2] = {&outRe, &outIm};
  for (int part = 0; part < 2; ++part) {
    if (!srcs[part]) continue;
    DeviceImages trimmed, full;
    HWS_STRY(trimmed.alloc(g, nImg * trimW * trimH));
    HWS_STRY(checkHip(g, hipMemcpyAsync(trimmed.p, srcs[part], nImg * trimW * trimH * sizeof(float),
                                        hipMemcpyHostToDevice, g->stream),
                      "model upload"));
    const float* cur = trimmed.p;
    if (trimW != imageW || trimH != imageH) {
      HWS_STRY(full.alloc(g, nImg * imageW * imageH));
      HWS_STRY(hws_untrim_device(g->cfg.device, g->stream, cur, trimW, trimH, full.p, imageW,
                                 imageH, nImg));
      cur = full.p;
    }
    HWS_STRY(outs[part]->alloc(g, nImg * size_t(g->W) * size_t(g->H)));
    if (imageW != size_t(g->W) || imageH != size_t(g->H)) {
      HWS_STRY(hws_resample_device(g->cfg.device, g->stream, cur, imageW, imageH, outs[part]->p,
                                   size_t(g->W), size_t(g->H), nImg));
    } else {
      HWS_STRY(checkHip(g, hipMemcpyAsync(outs[part]->p, cur,
                                          nImg * imageW * imageH * sizeof(float),
                                          hipMemcpyDeviceToDevice, g->stream),
                        "model copy"));
    }
    HWS_STRY(checkHip(g, hipStreamSynchronize(g->stream), "model preparation"));
  }
  return HWS_OK;
}

hws_status checkSizes(hws_gridder* g, size_t imageW, size_t imageH, size_t trimW, size_t trimH) {
  if (!validImageSize(imageW, imageH) || !validImageSize(trimW, trimH))
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "image and trimmed sizes must be even and positive");
  if (trimW > imageW || trimH > imageH)
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "trimmed size exceeds the image size");
  return HWS_OK;
}

}  // namespace

void freeStaging(hws_gridder* g) {
  for (int k = 0; k < 2; ++k) {
    if (g->pinnedFree[k]) (void)hipEventSynchronize(g->pinnedFree[k]);
    if (g->pinned[k]) (void)hipHostFree(g->pinned[k]);
    if (g->pinnedFree[k]) (void)hipEventDestroy(g->pinnedFree[k]);
    g->pinned[k] = nullptr;
    g->pinnedFree[k] = nullptr;
  }
  if (g->dStagedUvw) (void)hipFree(g->dStagedUvw);
  if (g->dStagedWl) (void)hipFree(g->dStagedWl);
  if (g->dStagedVis) (void)hipFree(g->dStagedVis);
  g->dStagedUvw = nullptr;
  g->dStagedWl = nullptr;
  g->dStagedVis = nullptr;
  g->stagedUvwCap = g->stagedWlCap = g->stagedVisCap = 0;
}

}  // namespace hws

using namespace hws;

extern "C" hws_status hws_stage_begin(hws_gridder* g, size_t max_rows, size_t nchan,
                                      const double* wavelengths, int with_visibilities) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (nchan == 0 || !wavelengths)
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "staging needs the channel wavelengths");
  if (max_rows > (size_t(1) << 30) / nchan)
    return sfail(g, HWS_ERR_UNSUPPORTED, "a plan holds at most 2^30 - 1 (row, channel) samples");
  (void)hipSetDevice(g->cfg.device);
  // a plan of the previously staged rows reads buffers that may move now
  if (g->plan.valid && g->plan.uvw == g->dStagedUvw) {
    g->plan.valid = false;
    ++g->generation->value;
  }
  HWS_STRY(grow(g, g->dStagedUvw, g->stagedUvwCap, max_rows * 3, "staged uvw"));
  HWS_STRY(grow(g, g->dStagedWl, g->stagedWlCap, nchan, "staged wavelengths"));
  if (with_visibilities)
    HWS_STRY(grow(g, g->dStagedVis, g->stagedVisCap, max_rows * nchan * size_t(g->nImg) * 2,
                  "staged visibilities"));
  HWS_STRY(checkHip(g, hipMemcpyAsync(g->dStagedWl, wavelengths, nchan * sizeof(double),
                                      hipMemcpyHostToDevice, g->stream),
                    "wavelength upload"));
  HWS_STRY(checkHip(g, hipStreamSynchronize(g->stream), "wavelength upload"));
  g->stagedMaxRows = max_rows;
  g->stagedNchan = nchan;
  g->stagedRows = 0;
  g->stagedWithVis = with_visibilities != 0;
  g->staging = true;
  g->staged = false;
  return HWS_OK;
}

extern "C" hws_status hws_stage_rows(hws_gridder* g, size_t nrow, const double* uvw_m,
                                     const float* vis) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->staging) return sfail(g, HWS_ERR_BAD_STATE, "hws_stage_begin has not been called");
  if (nrow == 0) return HWS_OK;
  if (!uvw_m || (g->stagedWithVis && !vis))
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "null row data");
  if (g->stagedRows + nrow > g->stagedMaxRows)
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "more rows staged than hws_stage_begin reserved");
  (void)hipSetDevice(g->cfg.device);
  HWS_STRY(upload(g, g->dStagedUvw + g->stagedRows * 3, uvw_m, nrow * 3 * sizeof(double)));
  if (g->stagedWithVis) {
    const size_t perRow = g->stagedNchan * size_t(g->nImg) * 2;
    HWS_STRY(upload(g, g->dStagedVis + g->stagedRows * perRow, vis,
                    nrow * perRow * sizeof(float)));
  }
  g->stagedRows += nrow;
  return HWS_OK;
}

extern "C" hws_status hws_stage_end(hws_gridder* g) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->staging) return sfail(g, HWS_ERR_BAD_STATE, "hws_stage_begin has not been called");
  (void)hipSetDevice(g->cfg.device);
  g->staging = false;
  HWS_STRY(hws_plan_device(g, g->stagedRows, g->stagedNchan, g->dStagedUvw, g->dStagedWl));
  g->staged = true;
  return HWS_OK;
}

extern "C" hws_status hws_invert_staged(hws_gridder* g, double multiplication_factor,
                                        size_t image_width, size_t image_height,
                                        size_t trim_width, size_t trim_height, float* real,
                                        float* imag) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->staged || !g->stagedWithVis)
    return sfail(g, HWS_ERR_BAD_STATE, "no staged rows with visibilities (hws_stage_end)");
  if (!real) return sfail(g, HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  HWS_STRY(checkSizes(g, image_width, image_height, trim_width, trim_height));
  (void)hipSetDevice(g->cfg.device);
  const size_t n = size_t(g->nImg) * size_t(g->W) * size_t(g->H);
  DeviceImages re, im;
  HWS_STRY(re.alloc(g, n));
  const bool wantImag = imag && g->anyComplex;
  if (wantImag) HWS_STRY(im.alloc(g, n));
  HWS_STRY(hws_invert_planned_device(g, g->dStagedVis, multiplication_factor, re.p,
                                     wantImag ? im.p : nullptr));
  return finishImages(g, re.p, wantImag ? im.p : nullptr, image_width, image_height, trim_width,
                      trim_height, real, wantImag ? imag : nullptr);
}

extern "C" hws_status hws_predict_staged(hws_gridder* g, const float* real, const float* imag,
                                         size_t trim_width, size_t trim_height,
                                         size_t image_width, size_t image_height,
                                         float* vis_out) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->staged) return sfail(g, HWS_ERR_BAD_STATE, "no staged rows (hws_stage_end)");
  if (!real || !vis_out) return sfail(g, HWS_ERR_INVALID_ARGUMENT, "null pointer");
  if (g->anyComplex && !imag)
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "Missing imaginary in complex prediction");
  HWS_STRY(checkSizes(g, image_width, image_height, trim_width, trim_height));
  (void)hipSetDevice(g->cfg.device);
  DeviceImages re, im, out;
  HWS_STRY(prepareModels(g, real, imag, trim_width, trim_height, image_width, image_height, re,
                         im));
  const size_t nvis = g->stagedRows * g->stagedNchan * size_t(g->nImg) * 2;
  HWS_STRY(out.alloc(g, nvis));
  HWS_STRY(hws_predict_planned_device(g, re.p, g->anyComplex ? im.p : nullptr, out.p));
  return download(g, vis_out, out.p, nvis * sizeof(float));
}

extern "C" hws_status hws_finalize_image_resampled(hws_gridder* g, double multiplication_factor,
                                                   size_t image_width, size_t image_height,
                                                   size_t trim_width, size_t trim_height,
                                                   float* real, float* imag) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!g->inversionActive) return sfail(g, HWS_ERR_BAD_STATE, "no inversion to finalise");
  if (!real) return sfail(g, HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  HWS_STRY(checkSizes(g, image_width, image_height, trim_width, trim_height));
  (void)hipSetDevice(g->cfg.device);
  const size_t n = size_t(g->nImg) * size_t(g->W) * size_t(g->H);
  DeviceImages re, im;
  HWS_STRY(re.alloc(g, n));
  const bool wantImag = imag && g->anyComplex;
  if (wantImag) HWS_STRY(im.alloc(g, n));
  HWS_STRY(finalizeImages(g, multiplication_factor, re.p, wantImag ? im.p : nullptr));
  g->inversionActive = false;
  return finishImages(g, re.p, wantImag ? im.p : nullptr, image_width, image_height, trim_width,
                      trim_height, real, wantImag ? imag : nullptr);
}

extern "C" hws_status hws_initialize_prediction_resampled(hws_gridder* g, const float* real,
                                                          const float* imag, size_t trim_width,
                                                          size_t trim_height, size_t image_width,
                                                          size_t image_height) {
  if (!g) return sfail(nullptr, HWS_ERR_INVALID_ARGUMENT, "null context");
  if (!real) return sfail(g, HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  if (g->anyComplex && !imag)
    return sfail(g, HWS_ERR_INVALID_ARGUMENT, "Missing imaginary in complex prediction");
  HWS_STRY(checkSizes(g, image_width, image_height, trim_width, trim_height));
  (void)hipSetDevice(g->cfg.device);
  DeviceImages re, im;
  HWS_STRY(prepareModels(g, real, imag, trim_width, trim_height, image_width, image_height, re,
                         im));
  // the context-size models, now in device memory, through the host-protocol
  // initialisation on its device buffers
  const size_t n = size_t(g->nImg) * size_t(g->W) * size_t(g->H);
  std::vector<float> hostRe(n), hostIm(g->anyComplex ? n : 0);
  HWS_STRY(checkHip(g, hipMemcpy(hostRe.data(), re.p, n * sizeof(float), hipMemcpyDeviceToHost),
                    "model download"));
  if (g->anyComplex)
    HWS_STRY(checkHip(g, hipMemcpy(hostIm.data(), im.p, n * sizeof(float),
                                   hipMemcpyDeviceToHost),
                      "model download"));
  return hws_initialize_prediction(g, hostRe.data(), g->anyComplex ? hostIm.data() : nullptr);
}
