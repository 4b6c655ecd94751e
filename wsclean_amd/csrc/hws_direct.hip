// FP64 direct Fourier transform on the GPU: the invert of
// DirectMSGridder<double> (gridding/directmsgridder.cpp:16-74, 76-114,
// 175-196) and its exact adjoint, the direct predict (the reference's
// DirectMSGridder::Predict throws, gridding/directmsgridder.cpp:167-171).
//
// Pixel (x, y) of a W x H image has l = (W/2 - x) dl + l_shift,
// m = (y - H/2) dm + m_shift and n' = sqrt(1 - l^2 - m^2) - 1 (0 outside the
// horizon). With phi_k(x, y) = u_k l + v_k m + w_k n' (uvw in wavelengths):
//   invert:  I(x, y) = f * sum_k Re(V_k exp(-2 pi i phi_k))
//                    = f * sum_k (Re V_k cos 2 pi phi_k + Im V_k sin 2 pi phi_k)
//   predict: V_k = sum_(x,y) I(x, y) exp(+2 pi i phi_k(x, y)).
// The reference evaluates the same sum per visibility into per-thread layers
// and remaps once at the end; here one thread owns one pixel (invert) or one
// visibility (predict) and sums in sample (pixel) order, so results are
// deterministic. The angle is evaluated with sincospi on 2 phi, whose
// argument reduction is exact; all arithmetic is FP64.
#include "hws_internal.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace hws {
namespace {

constexpr int kDirectThreads = 256;
constexpr int kDirectChunk = 256;

struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int device) {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = hipSetDevice(device) == hipSuccess;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

__device__ __forceinline__ void pixelLmn(int x, int y, int W, int H, double dl, double dm,
                                         double ls, double ms, double& l, double& m,
                                         double& n) {
  l = double(W / 2 - x) * dl + ls;
  m = double(y - H / 2) * dm + ms;
  const double r2 = l * l + m * m;
  n = r2 < 1.0 ? sqrt(1.0 - r2) - 1.0 : 0.0;
}

// One thread per pixel; the samples of a chunk are staged in LDS and read by
// every lane of the block (broadcast reads).
__global__ void __launch_bounds__(kDirectThreads)
directInvertKernel(const double* __restrict__ uvw, const double* __restrict__ wl, size_t nrow,
                   int nchan, const float2* __restrict__ vis, size_t stride, int W, int H,
                   double dl, double dm, double ls, double ms, double factor,
                   double* __restrict__ image) {
  __shared__ double sU[kDirectChunk], sV[kDirectChunk], sW[kDirectChunk];
  __shared__ double sRe[kDirectChunk], sIm[kDirectChunk];
  const size_t npix = size_t(W) * size_t(H);
  const size_t pix = size_t(blockIdx.x) * kDirectThreads + threadIdx.x;
  double l = 0.0, m = 0.0, n = 0.0;
  if (pix < npix) pixelLmn(int(pix % W), int(pix / W), W, H, dl, dm, ls, ms, l, m, n);
  const size_t ns = nrow * size_t(nchan);
  double acc = 0.0;
  for (size_t base = 0; base < ns; base += kDirectChunk) {
    const size_t i = base + threadIdx.x;
    if (i < ns) {
      const size_t row = i / size_t(nchan);
      const double lambda = wl[i - row * size_t(nchan)];
      sU[threadIdx.x] = uvw[row * 3] / lambda;
      sV[threadIdx.x] = uvw[row * 3 + 1] / lambda;
      sW[threadIdx.x] = uvw[row * 3 + 2] / lambda;
      const float2 v = vis[i * stride];
      sRe[threadIdx.x] = double(v.x);
      sIm[threadIdx.x] = double(v.y);
    }
    __syncthreads();
    const int cnt = int(min<size_t>(kDirectChunk, ns - base));
    for (int k = 0; k < cnt; ++k) {
      const double phi = sU[k] * l + sV[k] * m + sW[k] * n;
      double s, c;
      sincospi(2.0 * phi, &s, &c);
      acc += sRe[k] * c + sIm[k] * s;
    }
    __syncthreads();
  }
  if (pix < npix) image[pix] = acc * factor;
}

// Non-zero pixels of the model, in pixel order: count per block, scan, write.
__global__ void __launch_bounds__(kDirectThreads)
countNonzeroKernel(const float* __restrict__ image, size_t npix, uint32_t* __restrict__ counts) {
  const size_t pix = size_t(blockIdx.x) * kDirectThreads + threadIdx.x;
  const bool nz = pix < npix && image[pix] != 0.0f;
  const int c = __syncthreads_count(nz);
  if (threadIdx.x == 0) counts[blockIdx.x] = uint32_t(c);
}

// Exclusive scan of nblocks counts in one workgroup; total in counts[nblocks].
__global__ void __launch_bounds__(1024)
scanCountsKernel(uint32_t* __restrict__ counts, size_t nblocks) {
  __shared__ uint32_t part[1024];
  const size_t per = (nblocks + 1023) / 1024;
  const size_t b0 = threadIdx.x * per, b1 = min(nblocks, b0 + per);
  uint32_t s = 0;
  for (size_t b = b0; b < b1; ++b) s += counts[b];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (size_t b = b0; b < b1; ++b) {
    const uint32_t c = counts[b];
    counts[b] = run;
    run += c;
  }
  if (threadIdx.x == 1023) counts[nblocks] = part[1023];
}

struct PixelTerm {
  double l, m, n, value;
};

__global__ void __launch_bounds__(kDirectThreads)
compactNonzeroKernel(const float* __restrict__ image, int W, int H, double dl, double dm,
                     double ls, double ms, const uint32_t* __restrict__ offsets,
                     PixelTerm* __restrict__ terms) {
  __shared__ uint32_t waveBase[kDirectThreads / 64];
  const size_t npix = size_t(W) * size_t(H);
  const size_t pix = size_t(blockIdx.x) * kDirectThreads + threadIdx.x;
  const float value = pix < npix ? image[pix] : 0.0f;
  const bool nz = value != 0.0f;
  const unsigned long long mask = __ballot(nz);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) waveBase[wave] = uint32_t(__popcll(mask));
  __syncthreads();
  uint32_t base = offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) base += waveBase[w];
  if (nz) {
    const uint32_t rank = uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    PixelTerm t;
    pixelLmn(int(pix % W), int(pix / W), W, H, dl, dm, ls, ms, t.l, t.m, t.n);
    t.value = double(value);
    terms[base + rank] = t;
  }
}

// One thread per (row, channel) sample; the non-zero pixels of a chunk are
// staged in LDS.
__global__ void __launch_bounds__(kDirectThreads)
directPredictKernel(const double* __restrict__ uvw, const double* __restrict__ wl, size_t nrow,
                    int nchan, const PixelTerm* __restrict__ terms,
                    const uint32_t* __restrict__ nterms, float2* __restrict__ out,
                    size_t stride) {
  __shared__ double sL[kDirectChunk], sM[kDirectChunk], sN[kDirectChunk], sI[kDirectChunk];
  const size_t ns = nrow * size_t(nchan);
  const size_t i = size_t(blockIdx.x) * kDirectThreads + threadIdx.x;
  double u = 0.0, v = 0.0, w = 0.0;
  if (i < ns) {
    const size_t row = i / size_t(nchan);
    const double lambda = wl[i - row * size_t(nchan)];
    u = uvw[row * 3] / lambda;
    v = uvw[row * 3 + 1] / lambda;
    w = uvw[row * 3 + 2] / lambda;
  }
  const size_t nt = *nterms;
  double re = 0.0, im = 0.0;
  for (size_t base = 0; base < nt; base += kDirectChunk) {
    const size_t t = base + threadIdx.x;
    if (t < nt) {
      const PixelTerm p = terms[t];
      sL[threadIdx.x] = p.l;
      sM[threadIdx.x] = p.m;
      sN[threadIdx.x] = p.n;
      sI[threadIdx.x] = p.value;
    }
    __syncthreads();
    const int cnt = int(min<size_t>(kDirectChunk, nt - base));
    for (int k = 0; k < cnt; ++k) {
      const double phi = u * sL[k] + v * sM[k] + w * sN[k];
      double s, c;
      sincospi(2.0 * phi, &s, &c);
      re += sI[k] * c;
      im += sI[k] * s;
    }
    __syncthreads();
  }
  if (i < ns) out[i * stride] = make_float2(float(re), float(im));
}

hws_status validateGeometry(size_t width, size_t height, double dl, double dm, size_t nrow,
                            size_t nchan, size_t stride) {
  if (width == 0 || height == 0 || width > 65536 || height > 65536 ||
      width * height > (size_t(1) << 31))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "image size out of range");
  // WSClean rounds odd image sizes up to even ones (main/settings.cpp:358-367);
  // its pixel-to-(l, m) remap assumes even sizes
  if (width % 2 != 0 || height % 2 != 0)
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "image width and height must be even");
  if (!(dl > 0.0) || !(dm > 0.0))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "pixel sizes must be positive");
  if (nchan == 0 || nchan > (size_t(1) << 20) || stride == 0)
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "bad channel count or stride");
  // one thread per sample in one launch: the grid's work-item count must fit
  // in 32 bits
  if (nrow > (size_t(1) << 31) / nchan)
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT,
                        "too many samples for one launch (at most 2^31)");
  return HWS_OK;
}

// The direct predict's launches on `stream`: non-zero pixels of the model
// compacted into `terms` (room for every pixel) through the block counts in
// `counts` (pixBlocks + 1 entries, the last receives the total), then one
// thread per sample. Every entry of both buffers that a later kernel reads is
// written first, so neither needs clearing.
hws_status launchDirectPredict(hipStream_t stream, size_t width, size_t height, double dl,
                               double dm, double l_shift, double m_shift, const float* d_image,
                               size_t nrow, size_t nchan, const double* d_uvw_m,
                               const double* d_wavelengths, float* d_vis_out, size_t vis_stride,
                               uint32_t* counts, PixelTerm* terms) {
  const size_t npix = width * height;
  const size_t pixBlocks = (npix + kDirectThreads - 1) / kDirectThreads;
  countNonzeroKernel<<<unsigned(pixBlocks), kDirectThreads, 0, stream>>>(d_image, npix, counts);
  scanCountsKernel<<<1, 1024, 0, stream>>>(counts, pixBlocks);
  compactNonzeroKernel<<<unsigned(pixBlocks), kDirectThreads, 0, stream>>>(
      d_image, int(width), int(height), dl, dm, l_shift, m_shift, counts, terms);
  const size_t ns = nrow * nchan;
  if (ns)
    directPredictKernel<<<unsigned((ns + kDirectThreads - 1) / kDirectThreads), kDirectThreads,
                          0, stream>>>(d_uvw_m, d_wavelengths, nrow, int(nchan), terms,
                                       counts + pixBlocks, reinterpret_cast<float2*>(d_vis_out),
                                       vis_stride);
  return checkHip(nullptr, hipGetLastError(), "direct predict launch");
}

}  // namespace
}  // namespace hws

using namespace hws;

extern "C" hws_status hws_direct_invert_device(int device, void* hip_stream, size_t width,
                                               size_t height, double dl, double dm,
                                               double l_shift, double m_shift, size_t nrow,
                                               size_t nchan, const double* d_uvw_m,
                                               const double* d_wavelengths, const float* d_vis,
                                               size_t vis_stride, double multiplication_factor,
                                               double* d_image) {
  const hws_status valid = validateGeometry(width, height, dl, dm, nrow, nchan, vis_stride);
  if (valid != HWS_OK) return valid;
  if (!d_image || (nrow && (!d_uvw_m || !d_wavelengths || !d_vis)))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceGuard guard(device);
  if (!guard.ok) return setLastError(nullptr, HWS_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  const size_t npix = width * height;
  const unsigned blocks = unsigned((npix + kDirectThreads - 1) / kDirectThreads);
  directInvertKernel<<<blocks, kDirectThreads, 0, stream>>>(
      d_uvw_m, d_wavelengths, nrow, int(nchan), reinterpret_cast<const float2*>(d_vis),
      vis_stride, int(width), int(height), dl, dm, l_shift, m_shift, multiplication_factor,
      d_image);
  return checkHip(nullptr, hipGetLastError(), "direct invert launch");
}

extern "C" hws_status hws_direct_predict_device(int device, void* hip_stream, size_t width,
                                                size_t height, double dl, double dm,
                                                double l_shift, double m_shift,
                                                const float* d_image, size_t nrow, size_t nchan,
                                                const double* d_uvw_m,
                                                const double* d_wavelengths, float* d_vis_out,
                                                size_t vis_stride) {
  const hws_status valid = validateGeometry(width, height, dl, dm, nrow, nchan, vis_stride);
  if (valid != HWS_OK) return valid;
  if (!d_image || (nrow && (!d_uvw_m || !d_wavelengths || !d_vis_out)))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceGuard guard(device);
  if (!guard.ok) return setLastError(nullptr, HWS_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  const size_t npix = width * height;
  const size_t pixBlocks = (npix + kDirectThreads - 1) / kDirectThreads;
  uint32_t* counts = nullptr;
  PixelTerm* terms = nullptr;
  hws_status st = checkHip(nullptr, hipMalloc(&counts, (pixBlocks + 1) * sizeof(uint32_t)),
                           "direct predict counts");
  if (st == HWS_OK)
    st = checkHip(nullptr, hipMalloc(&terms, npix * sizeof(PixelTerm)), "direct predict terms");
  if (st == HWS_OK) {
    st = launchDirectPredict(stream, width, height, dl, dm, l_shift, m_shift, d_image, nrow,
                             nchan, d_uvw_m, d_wavelengths, d_vis_out, vis_stride, counts,
                             terms);
    // the scratch buffers are released once the transform has run
    const hws_status sy = checkHip(nullptr, hipStreamSynchronize(stream), "direct predict");
    if (st == HWS_OK) st = sy;
  }
  if (terms) (void)hipFree(terms);
  if (counts) (void)hipFree(counts);
  return st;
}

// ---- the direct predict without allocation -------------------------------
//
// The scratch of the direct predict (pixel terms and block counts) lives in
// one file-scope device array instead of hipMalloc'd buffers, for callers
// that cannot allocate (graph capture, serving loops). Its size is fixed:
// images of at most kNoAllocMaxPixels pixels. Each device holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (the array's name in host code
// is a host-side stand-in, not a device address).

namespace hws {
namespace {

constexpr size_t kNoAllocMaxPixels = size_t(1) << 20;
constexpr size_t kNoAllocMaxBlocks = kNoAllocMaxPixels / kDirectThreads;
constexpr size_t kScratchTermBytes = kNoAllocMaxPixels * sizeof(PixelTerm);
constexpr size_t kScratchBytes =
    kScratchTermBytes + (kNoAllocMaxBlocks + 1) * sizeof(uint32_t);  // 32 MiB + 16 KiB
constexpr int kScratchDevices = 64;

std::mutex scratchMutex;
void* scratchAddress[kScratchDevices] = {};

}  // namespace
}  // namespace hws

static __device__ __align__(256) unsigned char scratch[hws::kScratchBytes];

namespace hws {
namespace {

// The device address of `scratch` on the current device, looked up once per
// device (the caller has made `device` current).
hws_status scratchFor(int device, unsigned char** out) {
  if (device < 0 || device >= kScratchDevices)
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratchMutex);
  if (!scratchAddress[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return checkHip(nullptr, e, "scratch symbol address");
    if (!p) return setLastError(nullptr, HWS_ERR_HIP, "scratch symbol address is null");
    scratchAddress[device] = p;
  }
  *out = static_cast<unsigned char*>(scratchAddress[device]);
  return HWS_OK;
}

}  // namespace
}  // namespace hws

extern "C" hws_status hws_direct_predict_device_noalloc(
    int device, void* hip_stream, size_t width, size_t height, double dl, double dm,
    double l_shift, double m_shift, const float* d_image, size_t nrow, size_t nchan,
    const double* d_uvw_m, const double* d_wavelengths, float* d_vis_out, size_t vis_stride) {
  const hws_status valid = validateGeometry(width, height, dl, dm, nrow, nchan, vis_stride);
  if (valid != HWS_OK) return valid;
  if (width * height > kNoAllocMaxPixels)
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT,
                        "image too large for the no-allocation direct predict: " +
                            std::to_string(width * height) + " pixels, at most " +
                            std::to_string(kNoAllocMaxPixels));
  if (!d_image || (nrow && (!d_uvw_m || !d_wavelengths || !d_vis_out)))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceGuard guard(device);
  if (!guard.ok) return setLastError(nullptr, HWS_ERR_HIP, "hipSetDevice failed");
  unsigned char* base = nullptr;
  const hws_status st = scratchFor(device, &base);
  if (st != HWS_OK) return st;
  PixelTerm* terms = reinterpret_cast<PixelTerm*>(base);
  uint32_t* counts = reinterpret_cast<uint32_t*>(base + kScratchTermBytes);
  return launchDirectPredict(static_cast<hipStream_t>(hip_stream), width, height, dl, dm,
                             l_shift, m_shift, d_image, nrow, nchan, d_uvw_m, d_wavelengths,
                             d_vis_out, vis_stride, counts, terms);
}

// ---- host-pointer forms (for host callers such as the direct-FT plugin) ----

namespace hws {
namespace {

struct DeviceBuffers {
  void* p[5] = {};
  ~DeviceBuffers() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

hws_status devUpload(void** dst, const void* src, size_t bytes, hipStream_t stream) {
  if (hipMalloc(dst, bytes ? bytes : 1) != hipSuccess)
    return setLastError(nullptr, HWS_ERR_OUT_OF_MEMORY, "direct transform buffers");
  if (src && bytes)
    return checkHip(nullptr, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, stream),
                    "direct transform upload");
  return HWS_OK;
}

}  // namespace
}  // namespace hws

extern "C" hws_status hws_direct_invert(int device, size_t width, size_t height, double dl,
                                        double dm, double l_shift, double m_shift, size_t nrow,
                                        size_t nchan, const double* uvw_m,
                                        const double* wavelengths, const float* vis,
                                        double multiplication_factor, double* image) {
  const hws_status valid = validateGeometry(width, height, dl, dm, nrow, nchan, 1);
  if (valid != HWS_OK) return valid;
  if (!image || (nrow && (!uvw_m || !wavelengths || !vis)))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceGuard guard(device);
  if (!guard.ok) return setLastError(nullptr, HWS_ERR_HIP, "hipSetDevice failed");
  DeviceBuffers b;
  hipStream_t stream = nullptr;
  hws_status st = devUpload(&b.p[0], uvw_m, nrow * 3 * sizeof(double), stream);
  if (st == HWS_OK) st = devUpload(&b.p[1], wavelengths, nchan * sizeof(double), stream);
  if (st == HWS_OK) st = devUpload(&b.p[2], vis, nrow * nchan * 2 * sizeof(float), stream);
  if (st == HWS_OK) st = devUpload(&b.p[3], nullptr, width * height * sizeof(double), stream);
  if (st == HWS_OK)
    st = hws_direct_invert_device(device, stream, width, height, dl, dm, l_shift, m_shift, nrow,
                                  nchan, static_cast<const double*>(b.p[0]),
                                  static_cast<const double*>(b.p[1]),
                                  static_cast<const float*>(b.p[2]), 1, multiplication_factor,
                                  static_cast<double*>(b.p[3]));
  if (st == HWS_OK)
    st = checkHip(nullptr, hipMemcpy(image, b.p[3], width * height * sizeof(double),
                                     hipMemcpyDeviceToHost),
                  "direct transform download");
  return st;
}

extern "C" hws_status hws_direct_predict(int device, size_t width, size_t height, double dl,
                                         double dm, double l_shift, double m_shift,
                                         const float* image, size_t nrow, size_t nchan,
                                         const double* uvw_m, const double* wavelengths,
                                         float* vis_out) {
  const hws_status valid = validateGeometry(width, height, dl, dm, nrow, nchan, 1);
  if (valid != HWS_OK) return valid;
  if (!image || (nrow && (!uvw_m || !wavelengths || !vis_out)))
    return setLastError(nullptr, HWS_ERR_INVALID_ARGUMENT, "null data pointer");
  DeviceGuard guard(device);
  if (!guard.ok) return setLastError(nullptr, HWS_ERR_HIP, "hipSetDevice failed");
  DeviceBuffers b;
  hipStream_t stream = nullptr;
  hws_status st = devUpload(&b.p[0], uvw_m, nrow * 3 * sizeof(double), stream);
  if (st == HWS_OK) st = devUpload(&b.p[1], wavelengths, nchan * sizeof(double), stream);
  if (st == HWS_OK) st = devUpload(&b.p[2], image, width * height * sizeof(float), stream);
  if (st == HWS_OK) st = devUpload(&b.p[3], nullptr, nrow * nchan * 2 * sizeof(float), stream);
  if (st == HWS_OK)
    st = hws_direct_predict_device(device, stream, width, height, dl, dm, l_shift, m_shift,
                                   static_cast<const float*>(b.p[2]), nrow, nchan,
                                   static_cast<const double*>(b.p[0]),
                                   static_cast<const double*>(b.p[1]),
                                   static_cast<float*>(b.p[3]), 1);
  if (st == HWS_OK)
    st = checkHip(nullptr, hipMemcpy(vis_out, b.p[3], nrow * nchan * 2 * sizeof(float),
                                     hipMemcpyDeviceToHost),
                  "direct transform download");
  return st;
}
