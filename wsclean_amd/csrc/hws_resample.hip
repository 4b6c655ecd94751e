// Small inversion on the GPU: the FFT resampler WSMSGridder applies between
// the inversion size and the image size (schaapcommon::fft::Resampler,
// gridding/wsmsgridder.cpp:452-474 and 521-551; schaapcommon is not vendored
// in the reference tree), and the centre-keeping image trim and untrim
// (aocommon::Image::Trim/Untrim, gridding/wsmsgridder.cpp:476-484, 516-520).
//
// The resampler's published algorithm: a real-to-complex transform of the
// in_w x in_h image, the low-frequency block both sizes share copied into an
// out_w x out_h half spectrum (columns 0 .. min(in_w, out_w)/2; rows
// 0 .. m - 1 and the last m rows, m = min(in_h, out_h)/2), a complex-to-real
// transform, and a factor 1 / (in_w in_h). Every step is linear and acts on
// the two axes separately, so
//   out = Re(My in Nx^T) / (in_w in_h)
// with, for t = (oy / out_h - jy / in_h) and s = (ox / out_w - jx / in_w),
//   My[oy][jy] = sum_{f=-m}^{m-1} exp(2 pi i f t)
//              = exp(-i pi t) sin(2 pi m t) / sin(pi t)   (2m at t = 0),
//   Nx[ox][jx] = sum_{k=0}^{K-1} c_k exp(2 pi i k s),  K = min(in_w, out_w)/2 + 1,
// c_0 = 1, c_k = 2 (the Hermitian partner the complex-to-real transform
// adds), c_{K-1} = 1 when column K-1 is the output's Nyquist column; with
// S(s, K) = sum_{k<K} exp(2 pi i k s) = exp(i pi (K-1) s) sin(pi K s) /
// sin(pi s), Nx = 2 S(s, K) - 1 - [Nyquist] exp(2 pi i (K-1) s).
// The matrices are evaluated in FP64 from these closed forms (s and t as
// exact integer ratios) and applied with FP64 matrix products (four DGEMMs
// per image), which keeps the result within FP64 rounding of the transform
// the FFTs compute.
#include "hws_internal.h"

#include <rocblas/rocblas.h>

#include <mutex>
#include <string>

namespace hws {
namespace {

constexpr int kRThreads = 256;

#define HWS_RTRY(expr)             \
  do {                             \
    const hws_status st_ = (expr); \
    if (st_ != HWS_OK) return st_; \
  } while (0)

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

// My (rows = out_h, cols = in_h), row-major, real and imaginary parts.
__global__ void __launch_bounds__(kRThreads)
rowMatrixKernel(int outH, int inH, int m, double* __restrict__ re, double* __restrict__ im) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i >= size_t(outH) * size_t(inH)) return;
  const long long oy = (long long)(i / size_t(inH)), jy = (long long)(i % size_t(inH));
  const long long num = oy * inH - jy * outH;  // t = num / (out_h in_h), exact
  double vr, vi;
  if (num == 0) {
    vr = 2.0 * m;
    vi = 0.0;
  } else {
    const double t = double(num) / (double(outH) * double(inH));
    double s1, c1;
    sincospi(t, &s1, &c1);
    const double ratio = sinpi(2.0 * m * t) / s1;
    vr = c1 * ratio;   // exp(-i pi t) * ratio
    vi = -s1 * ratio;
  }
  re[i] = vr;
  im[i] = vi;
}

// Nx (rows = out_w, cols = in_w), row-major.
__global__ void __launch_bounds__(kRThreads)
colMatrixKernel(int outW, int inW, int K, int nyquist, double* __restrict__ re,
                double* __restrict__ im) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i >= size_t(outW) * size_t(inW)) return;
  const long long ox = (long long)(i / size_t(inW)), jx = (long long)(i % size_t(inW));
  const long long num = ox * inW - jx * outW;  // s = num / (out_w in_w), exact
  double sr, si;
  double er = 1.0, ei = 0.0;  // exp(2 pi i (K - 1) s)
  if (num == 0) {
    sr = K;
    si = 0.0;
  } else {
    const double s = double(num) / (double(outW) * double(inW));
    double s1, c1, sh, ch;
    sincospi(s, &s1, &c1);
    sincospi(double(K - 1) * s, &sh, &ch);
    const double ratio = sinpi(double(K) * s) / s1;
    sr = ch * ratio;
    si = sh * ratio;
    sincospi(2.0 * double(K - 1) * s, &ei, &er);
  }
  double vr = 2.0 * sr - 1.0, vi = 2.0 * si;
  if (nyquist) {
    vr -= er;
    vi -= ei;
  }
  re[i] = vr;
  im[i] = vi;
}

__global__ void __launch_bounds__(kRThreads)
toDoubleKernel(const float* __restrict__ in, double* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i < n) out[i] = double(in[i]);
}

__global__ void __launch_bounds__(kRThreads)
toFloatKernel(const double* __restrict__ in, float* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i < n) out[i] = float(in[i]);
}

// out[y][x] = in[y + sy][x + sx] (trim) for the th x tw window, per image.
__global__ void __launch_bounds__(kRThreads)
trimKernel(const float* __restrict__ in, size_t w, size_t h, float* __restrict__ out,
           size_t tw, size_t th) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i >= tw * th) return;
  const size_t img = blockIdx.y;
  const size_t y = i / tw, x = i % tw;
  const size_t sx = (w - tw) / 2, sy = (h - th) / 2;
  out[img * tw * th + i] = in[img * w * h + (y + sy) * w + (x + sx)];
}

// out = zeros(w x h) with in placed at (sx, sy) (untrim), per image.
__global__ void __launch_bounds__(kRThreads)
untrimKernel(const float* __restrict__ in, size_t tw, size_t th, float* __restrict__ out,
             size_t w, size_t h) {
  const size_t i = size_t(blockIdx.x) * kRThreads + threadIdx.x;
  if (i >= w * h) return;
  const size_t img = blockIdx.y;
  const size_t y = i / w, x = i % w;
  const size_t sx = (w - tw) / 2, sy = (h - th) / 2;
  float v = 0.0f;
  if (x >= sx && x < sx + tw && y >= sy && y < sy + th)
    v = in[img * tw * th + (y - sy) * tw + (x - sx)];
  out[img * w * h + i] = v;
}

hws_status fail(hws_status s, const std::string& m) { return setLastError(nullptr, s, m); }

std::mutex blasMutex;
rocblas_handle blasHandle[64] = {};

hws_status blasFor(int device, rocblas_handle* out) {
  if (device < 0 || device >= 64) return fail(HWS_ERR_INVALID_ARGUMENT, "device out of range");
  if (!blasHandle[device] && rocblas_create_handle(&blasHandle[device]) != rocblas_status_success)
    return fail(HWS_ERR_HIP, "rocblas_create_handle failed");
  *out = blasHandle[device];
  return HWS_OK;
}

hws_status checkBlas(rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return HWS_OK;
  return fail(HWS_ERR_HIP, std::string(what) + ": " + rocblas_status_to_string(s));
}

hws_status checkSizes(size_t w, size_t h, const char* what) {
  if (w == 0 || h == 0 || w % 2 != 0 || h % 2 != 0 || w > 65536 || h > 65536)
    return fail(HWS_ERR_INVALID_ARGUMENT, std::string(what) + ": sizes must be even, 2..65536");
  return HWS_OK;
}

unsigned blocksOf(size_t n) { return unsigned((n + kRThreads - 1) / kRThreads); }

struct Scratch {
  double* p[8] = {};
  ~Scratch() {
    for (double* q : p)
      if (q) (void)hipFree(q);
  }
};

}  // namespace
}  // namespace hws

using namespace hws;

extern "C" hws_status hws_resample_device(int device, void* hip_stream, const float* d_in,
                                          size_t in_w, size_t in_h, float* d_out, size_t out_w,
                                          size_t out_h, size_t n_images) {
  HWS_RTRY(checkSizes(in_w, in_h, "resample input"));
  HWS_RTRY(checkSizes(out_w, out_h, "resample output"));
  if (n_images == 0) return HWS_OK;
  if (!d_in || !d_out) return fail(HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  if (n_images > 65535) return fail(HWS_ERR_INVALID_ARGUMENT, "too many images");
  DeviceScope scope(device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  const size_t K = std::min(in_w / 2 + 1, out_w / 2 + 1);
  const int nyquist = (2 * (K - 1) == out_w) ? 1 : 0;
  const size_t m = std::min(in_h, out_h) / 2;
  Scratch sc;
  // 0,1: My re/im (out_h x in_h); 2,3: Nx re/im (out_w x in_w);
  // 4,5: T re/im (in_h x out_w); 6: input (in_h x in_w); 7: output
  const size_t sizes[8] = {out_h * in_h, out_h * in_h, out_w * in_w, out_w * in_w,
                           in_h * out_w, in_h * out_w, in_h * in_w, out_h * out_w};
  for (int i = 0; i < 8; ++i)
    if (hipMalloc(&sc.p[i], sizes[i] * sizeof(double)) != hipSuccess)
      return fail(HWS_ERR_OUT_OF_MEMORY, "resampler buffers");
  hipLaunchKernelGGL(rowMatrixKernel, dim3(blocksOf(sizes[0])), dim3(kRThreads), 0, stream,
                     int(out_h), int(in_h), int(m), sc.p[0], sc.p[1]);
  hipLaunchKernelGGL(colMatrixKernel, dim3(blocksOf(sizes[2])), dim3(kRThreads), 0, stream,
                     int(out_w), int(in_w), int(K), nyquist, sc.p[2], sc.p[3]);
  HWS_RTRY(checkHip(nullptr, hipGetLastError(), "resampler matrices"));
  std::lock_guard<std::mutex> lock(blasMutex);
  rocblas_handle h = nullptr;
  HWS_RTRY(blasFor(device, &h));
  HWS_RTRY(checkBlas(rocblas_set_stream(h, stream), "rocblas_set_stream"));
  const double factor = 1.0 / (double(in_w) * double(in_h));
  const double one = 1.0, zero = 0.0, negFactor = -factor;
  for (size_t img = 0; img < n_images; ++img) {
    hipLaunchKernelGGL(toDoubleKernel, dim3(blocksOf(sizes[6])), dim3(kRThreads), 0, stream,
                       d_in + img * in_w * in_h, sc.p[6], sizes[6]);
    // T = in Nx^T (row-major): column-major T^T (out_w x in_h) = Nx (out_w x
    // in_w) * in^T (in_w x in_h), Nx stored row-major = its transpose column-major
    for (int part = 0; part < 2; ++part)
      HWS_RTRY(checkBlas(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none,
                                       rocblas_int(out_w), rocblas_int(in_h), rocblas_int(in_w),
                                       &one, sc.p[2 + part], rocblas_int(in_w), sc.p[6],
                                       rocblas_int(in_w), &zero, sc.p[4 + part],
                                       rocblas_int(out_w)),
                         "resampler column product"));
    // out = (My_re T_re - My_im T_im) * factor (row-major): column-major
    // out^T (out_w x out_h) = T^T (out_w x in_h) * My^T (in_h x out_h)
    HWS_RTRY(checkBlas(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none,
                                     rocblas_int(out_w), rocblas_int(out_h), rocblas_int(in_h),
                                     &factor, sc.p[4], rocblas_int(out_w), sc.p[0],
                                     rocblas_int(in_h), &zero, sc.p[7], rocblas_int(out_w)),
                       "resampler row product"));
    HWS_RTRY(checkBlas(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none,
                                     rocblas_int(out_w), rocblas_int(out_h), rocblas_int(in_h),
                                     &negFactor, sc.p[5], rocblas_int(out_w), sc.p[1],
                                     rocblas_int(in_h), &one, sc.p[7], rocblas_int(out_w)),
                       "resampler row product"));
    hipLaunchKernelGGL(toFloatKernel, dim3(blocksOf(sizes[7])), dim3(kRThreads), 0, stream,
                       sc.p[7], d_out + img * out_w * out_h, sizes[7]);
    HWS_RTRY(checkHip(nullptr, hipGetLastError(), "resampler conversion"));
  }
  // the scratch buffers are freed when this call returns
  return checkHip(nullptr, hipStreamSynchronize(stream), "resample");
}

extern "C" hws_status hws_trim_device(int device, void* hip_stream, const float* d_in, size_t width,
                                      size_t height, float* d_out, size_t trim_width,
                                      size_t trim_height, size_t n_images) {
  HWS_RTRY(checkSizes(width, height, "trim input"));
  HWS_RTRY(checkSizes(trim_width, trim_height, "trim output"));
  if (trim_width > width || trim_height > height)
    return fail(HWS_ERR_INVALID_ARGUMENT, "trimmed size exceeds the image");
  if (n_images == 0) return HWS_OK;
  if (!d_in || !d_out) return fail(HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  if (n_images > 65535) return fail(HWS_ERR_INVALID_ARGUMENT, "too many images");
  DeviceScope scope(device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  hipLaunchKernelGGL(trimKernel, dim3(blocksOf(trim_width * trim_height), unsigned(n_images)),
                     dim3(kRThreads), 0, static_cast<hipStream_t>(hip_stream), d_in, width,
                     height, d_out, trim_width, trim_height);
  return checkHip(nullptr, hipGetLastError(), "trim launch");
}

extern "C" hws_status hws_untrim_device(int device, void* hip_stream, const float* d_in,
                                        size_t trim_width, size_t trim_height, float* d_out,
                                        size_t width, size_t height, size_t n_images) {
  HWS_RTRY(checkSizes(width, height, "untrim output"));
  HWS_RTRY(checkSizes(trim_width, trim_height, "untrim input"));
  if (trim_width > width || trim_height > height)
    return fail(HWS_ERR_INVALID_ARGUMENT, "trimmed size exceeds the image");
  if (n_images == 0) return HWS_OK;
  if (!d_in || !d_out) return fail(HWS_ERR_INVALID_ARGUMENT, "null image pointer");
  if (n_images > 65535) return fail(HWS_ERR_INVALID_ARGUMENT, "too many images");
  DeviceScope scope(device);
  if (!scope.ok) return fail(HWS_ERR_HIP, "hipSetDevice failed");
  hipLaunchKernelGGL(untrimKernel, dim3(blocksOf(width * height), unsigned(n_images)),
                     dim3(kRThreads), 0, static_cast<hipStream_t>(hip_stream), d_in, trim_width,
                     trim_height, d_out, width, height);
  return checkHip(nullptr, hipGetLastError(), "untrim launch");
}
