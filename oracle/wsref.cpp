// ORACLE / TEST INFRASTRUCTURE ONLY (see oracle.h).
// CPU restatement of WStackingGridder<T>; each function cites the reference
// lines (gridding/wstackinggridder.{h,cpp}) whose arithmetic it follows.
#include "wsref.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;

// gridding/wstackinggridder.cpp:425-447
void makeKaiserBessel(std::vector<double>& kernel, double alpha,
                      size_t overSampling, bool withSinc) {
  const size_t n = kernel.size(), mid = n / 2;
  std::vector<double> sincKernel(mid + 1);
  const double filterRatio = 1.0 / double(overSampling);
  sincKernel[0] = filterRatio;
  for (size_t i = 1; i != mid + 1; i++) {
    const double x = i;
    sincKernel[i] =
        withSinc ? (std::sin(kPi * filterRatio * x) / (kPi * x)) : filterRatio;
  }
  const double normFactor = double(overSampling) / Bessel0(alpha, 1e-8);
  for (size_t i = 0; i != mid + 1; i++) {
    const double term = double(i) / mid;
    kernel[mid + i] = sincKernel[i] *
                      Bessel0(alpha * std::sqrt(1.0 - (term * term)), 1e-8) *
                      normFactor;
  }
  for (size_t i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
}

// gridding/wstackinggridder.cpp:449-462
void makeRectangular(std::vector<double>& kernel, size_t overSampling) {
  const size_t n = kernel.size(), mid = n / 2;
  const double filterRatio = 1.0 / double(overSampling);
  kernel[mid] = 1.0;
  const double normFactor = double(overSampling);
  for (size_t i = 1; i != mid + 1; i++) {
    const double x = i;
    kernel[mid + i] = normFactor * std::sin(kPi * filterRatio * x) / (kPi * x);
  }
  for (size_t i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
}

// gridding/wstackinggridder.cpp:464-479
void makeGaussian(std::vector<double>& kernel, size_t overSampling,
                  bool withSinc) {
  const size_t n = kernel.size(), mid = n / 2;
  const double filterRatio = 1.0 / double(overSampling);
  kernel[mid] = 1.0;
  const double normFactor = double(overSampling);
  for (size_t i = 1; i != mid + 1; i++) {
    const double x = i, y = x * x * 3.0 * 3.0 / (mid * mid);
    kernel[mid + i] = withSinc ? normFactor * std::sin(kPi * filterRatio * x) /
                                     (kPi * x) * std::exp(y / -2.0)
                               : std::exp(y / -2.0);
  }
  for (size_t i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
}

// gridding/wstackinggridder.cpp:481-502
void makeBlackmanNuttall(std::vector<double>& kernel, size_t overSampling,
                         bool withSinc) {
  const size_t n = kernel.size(), mid = n / 2;
  const double filterRatio = 1.0 / double(overSampling);
  kernel[mid] = 1.0;
  const double normFactor = double(overSampling);
  const double a0 = 0.3635819, a1 = 0.4891775, a2 = 0.1365995, a3 = 0.0106411;
  for (size_t i = 1; i != mid + 1; i++) {
    const double x = i, y = mid + i;
    const double bn = a0 - a1 * std::cos(2.0 * kPi * y / (n - 1)) +
                      a2 * std::cos(4.0 * kPi * y / (n - 1)) -
                      a3 * std::cos(6.0 * kPi * y / (n - 1));
    if (withSinc)
      kernel[mid + i] =
          normFactor * std::sin(kPi * filterRatio * x) / (kPi * x) * bn;
    else
      kernel[mid + i] = bn;
  }
  for (size_t i = 0; i != mid; i++) kernel[i] = kernel[n - 1 - i];
}

bool fillKernel(KernelMode mode, std::vector<double>& v, size_t os,
                bool nnIsDelta) {
  const double alpha = 8.6;  // gridding/wstackinggridder.cpp:342, 391
  switch (mode) {
    case KernelMode::NearestNeighbour:
      if (nnIsDelta) {
        std::fill(v.begin(), v.end(), 0.0);
        v[v.size() / 2] = 1.0;
      } else {
        makeKaiserBessel(v, alpha, os, true);
      }
      return true;
    case KernelMode::KaiserBessel:
      makeKaiserBessel(v, alpha, os, true);
      return true;
    case KernelMode::KaiserBesselWithoutSinc:
      makeKaiserBessel(v, alpha, os, false);
      return true;
    case KernelMode::Rectangular:
      makeRectangular(v, os);
      return true;
    case KernelMode::Gaussian:
      makeGaussian(v, os, true);
      return true;
    case KernelMode::GaussianWithoutSinc:
      makeGaussian(v, os, false);
      return true;
    case KernelMode::BlackmanNuttall:
      makeBlackmanNuttall(v, os, true);
      return true;
    case KernelMode::BlackmanNuttallWithoutSinc:
      makeBlackmanNuttall(v, os, false);
      return true;
    case KernelMode::BlackmanHarris:
      return false;
  }
  return false;
}
}  // namespace

double Bessel0(double x, double precision) {
  double d = 0.0, ds = 1.0, sum = 1.0;
  do {
    d += 2.0;
    ds *= x * x / (d * d);
    sum += ds;
  } while (ds > sum * precision);
  return sum;
}

bool MakeKernel1D(KernelMode mode, double* kernel, size_t overSampling,
                  size_t size) {
  std::vector<double> v(overSampling * size, 0.0);
  if (!fillKernel(mode, v, overSampling, true)) return false;
  std::copy(v.begin(), v.end(), kernel);
  return true;
}

template <typename T>
WStackingRef<T>::WStackingRef(size_t width, size_t height, double pixelSizeX,
                              double pixelSizeY, size_t fftThreadCount,
                              size_t kernelSize, size_t overSamplingFactor)
    : width_(width),
      height_(height),
      pixelSizeX_(pixelSizeX),
      pixelSizeY_(pixelSizeY),
      overSampling_(overSamplingFactor),
      kernelSize_(kernelSize),
      image_(fftThreadCount),
      imageImag_(fftThreadCount),
      nFFTThreads_(fftThreadCount) {
  makeKernels();
}

template <typename T>
void WStackingRef<T>::SetGridMode(KernelMode mode) {
  if (mode != gridMode_) {
    gridMode_ = mode;
    if (gridMode_ != KernelMode::NearestNeighbour) makeKernels();
  }
}

// gridding/wstackinggridder.cpp:338-385
template <typename T>
void WStackingRef<T>::makeKernels() {
  kernels_.assign(overSampling_, std::vector<T>());
  k1d_.assign(kernelSize_ * overSampling_, 0.0);
  if (!fillKernel(gridMode_, k1d_, overSampling_, false))
    throw std::runtime_error(
        "Blackman-Harris kernel not supported by W-Stacking gridder");
  for (size_t i = 0; i != overSampling_; ++i) {
    std::vector<T>& kernel = kernels_[overSampling_ - i - 1];
    kernel.resize(kernelSize_);
    for (size_t x = 0; x != kernelSize_; ++x)
      kernel[x] = T(k1d_[x * overSampling_ + i]);
  }
}

// gridding/wstackinggridder.cpp:63-116
template <typename T>
void WStackingRef<T>::PrepareWLayers(size_t nWLayers, double maxMem,
                                     double minW, double maxW) {
  minW_ = minW;
  maxW_ = maxW;
  nW_ = nWLayers;
  if (minW_ == maxW_) maxW_ += 1.0;

  const size_t nrCopies = std::min<size_t>(nFFTThreads_, nW_);
  const double memPerImage = double(width_ * height_ * sizeof(T));
  const double memPerCore = memPerImage * 5.0;
  double remainingMem = maxMem - nrCopies * memPerCore;
  if (remainingMem <= memPerImage * nFFTThreads_) {
    nFFTThreads_ = size_t(maxMem * 3.0 / (5.0 * memPerCore));
    if (nFFTThreads_ == 0) nFFTThreads_ = 1;
    remainingMem = maxMem - nFFTThreads_ * memPerCore;
  }
  if (image_.size() < nFFTThreads_) {
    image_.resize(nFFTThreads_);
    imageImag_.resize(nFFTThreads_);
  }
  for (size_t i = 0; i != nFFTThreads_; ++i) {
    image_[i].assign(width_ * height_, T(0));
    if (isComplex_) imageImag_[i].assign(width_ * height_, T(0));
  }
  int maxNWLayersPerPass = int(remainingMem / (2.0 * memPerImage));
  if (maxNWLayersPerPass < 1) maxNWLayersPerPass = 1;
  nPasses_ = (nWLayers + maxNWLayersPerPass - 1) / maxNWLayersPerPass;
  if (nPasses_ == 0) nPasses_ = 1;
  pass_ = 0;
}

template <typename T>
size_t WStackingRef<T>::WToLayer(double w) const {
  if (nW_ == 1) return 0;
  double r;
  if (isComplex_)
    r = std::round((w + maxW_) * (nW_ - 1) / (maxW_ + maxW_));
  else
    r = std::round((std::fabs(w) - minW_) * (nW_ - 1) / (maxW_ - minW_));
  if (r < 0.0) return std::numeric_limits<size_t>::max();
  return size_t(r);
}

template <typename T>
double WStackingRef<T>::LayerToW(size_t layer) const {
  if (nW_ == 1) return 0.0;
  if (isComplex_) return layer * (maxW_ + maxW_) / (nW_ - 1) - maxW_;
  return layer * (maxW_ - minW_) / (nW_ - 1) + minW_;
}

template <typename T>
bool WStackingRef<T>::IsInLayerRange(double w) const {
  const size_t layer = WToLayer(w);
  return layer >= LayerRangeStart(pass_) && layer < LayerRangeStart(pass_ + 1);
}

template <typename T>
bool WStackingRef<T>::IsInLayerRange(double wStart, double wEnd) const {
  const size_t rangeStart = LayerRangeStart(pass_),
               rangeEnd = LayerRangeStart(pass_ + 1), l1 = WToLayer(wStart);
  if (l1 >= rangeStart && l1 < rangeEnd) return true;
  const size_t l2 = WToLayer(wEnd);
  return (l2 >= rangeStart && l2 < rangeEnd) ||
         (l1 < rangeStart && l2 >= rangeEnd) ||
         (l2 < rangeStart && l1 >= rangeEnd);
}

// gridding/wstackinggridder.cpp:851-872
template <typename T>
void WStackingRef<T>::initializeSqrtLMLookupTable() {
  sqrtLM_.resize(width_ * height_);
  T* iter = sqrtLM_.data();
  for (size_t y = 0; y != height_; ++y) {
    size_t ySrc = (height_ - y) + height_ / 2;
    if (ySrc >= height_) ySrc -= height_;
    const double m = ((double)ySrc - (height_ / 2)) * pixelSizeY_ + mShift_;
    for (size_t x = 0; x != width_; ++x) {
      size_t xSrc = x + width_ / 2;
      if (xSrc >= width_) xSrc -= width_;
      const double l = ((width_ / 2) - (double)xSrc) * pixelSizeX_ + lShift_;
      *iter++ = (l * l + m * m < 1.0) ? T(std::sqrt(1.0 - l * l - m * m) - 1.0)
                                      : T(0.0);
    }
  }
}

// gridding/wstackinggridder.cpp:909-932
template <typename T>
void WStackingRef<T>::initializeSqrtLMLookupTableForSampling() {
  sqrtLM_.resize(width_ * height_);
  T* iter = sqrtLM_.data();
  for (size_t y = 0; y != height_; ++y) {
    size_t yDest = y + height_ / 2;
    if (yDest >= height_) yDest -= height_;
    const double m = ((double)yDest - (height_ / 2)) * pixelSizeY_ + mShift_;
    for (size_t x = 0; x != width_; ++x) {
      size_t xDest = (width_ - x) + width_ / 2;
      if (xDest >= width_) xDest -= width_;
      const double l = ((width_ / 2) - (double)xDest) * pixelSizeX_ + lShift_;
      *iter++ = (l * l + m * m < 1.0) ? T(std::sqrt(1.0 - l * l - m * m) - 1.0)
                                      : T(0.0);
    }
  }
}

// gridding/wstackinggridder.cpp:127-137 (with initializeLayeredUVData :118-125)
template <typename T>
void WStackingRef<T>::StartInversionPass(size_t pass) {
  initializeSqrtLMLookupTable();
  pass_ = pass;
  const size_t n = LayerRangeStart(pass + 1) - LayerRangeStart(pass);
  layers_.resize(n);
  for (auto& l : layers_) l.assign(width_ * height_, std::complex<T>(0, 0));
}

// gridding/wstackinggridder.cpp:517-597
template <typename T>
void WStackingRef<T>::AddDataSample(std::complex<float> sample, double u,
                                    double v, double w) {
  const size_t layerOffset = LayerRangeStart(pass_),
               layerRangeEnd = LayerRangeStart(pass_ + 1);
  if (w < 0.0 && !isComplex_) {
    u = -u;
    v = -v;
    w = -w;
    sample = std::conj(sample);
  }
  const size_t wLayer = WToLayer(w);
  if (!(wLayer >= layerOffset && wLayer < layerRangeEnd)) return;
  std::complex<T>* uvData = layers_[wLayer - layerOffset].data();
  const int W = int(width_), H = int(height_);
  if (gridMode_ == KernelMode::NearestNeighbour) {
    int x = int(std::round(u * pixelSizeX_ * width_)),
        y = int(std::round(v * pixelSizeY_ * height_));
    if (x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2) {
      if (x < 0) x += W;
      if (y < 0) y += H;
      uvData[x + y * width_] += std::complex<T>(sample.real(), sample.imag());
    }
    return;
  }
  const double xExact = u * pixelSizeX_ * width_,
               yExact = v * pixelSizeY_ * height_;
  int x = int(std::round(xExact));
  int y = int(std::round(yExact));
  const int os = int(overSampling_);
  int xKernelIndex = int(std::round((xExact - double(x)) * overSampling_));
  int yKernelIndex = int(std::round((yExact - double(y)) * overSampling_));
  xKernelIndex = (xKernelIndex + (os * 3) / 2) % os;
  yKernelIndex = (yKernelIndex + (os * 3) / 2) % os;
  const std::vector<T>& xKernel = kernels_[xKernelIndex];
  const std::vector<T>& yKernel = kernels_[yKernelIndex];
  const int mid = int(kernelSize_) / 2;
  if (x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2) {
    if (x < 0) x += W;
    if (y < 0) y += H;
    for (size_t j = 0; j != kernelSize_; ++j) {
      const T yKernelValue = yKernel[j];
      const size_t cy = ((y + j + height_ - mid) % height_) * width_;
      for (size_t i = 0; i != kernelSize_; ++i) {
        const size_t cx = (x + i + width_ - mid) % width_;
        const T kernelValue = yKernelValue * xKernel[i];
        uvData[cx + cy] += std::complex<T>(sample.real() * kernelValue,
                                           sample.imag() * kernelValue);
      }
    }
  }
}

// gridding/wstackinggridder.cpp:874-907
template <typename T>
template <bool IsComplexImpl>
void WStackingRef<T>::projectOnImageAndCorrect(const std::complex<T>* source,
                                               double w, size_t threadIndex) {
  T* dataReal = image_[threadIndex].data();
  T* dataImaginary = IsComplexImpl ? imageImag_[threadIndex].data() : nullptr;
  const double twoPiW = -2.0 * kPi * w;
  const T* sqrtLMIter = sqrtLM_.data();
  for (size_t y = 0; y != height_; ++y) {
    size_t ySrc = (height_ - y) + height_ / 2;
    if (ySrc >= height_) ySrc -= height_;
    for (size_t x = 0; x != width_; ++x) {
      size_t xSrc = x + width_ / 2;
      if (xSrc >= width_) xSrc -= width_;
      const double rad = twoPiW * *sqrtLMIter, s = std::sin(rad),
                   c = std::cos(rad);
      dataReal[xSrc + ySrc * width_] += source->real() * c - source->imag() * s;
      if (IsComplexImpl)
        dataImaginary[xSrc + ySrc * width_] +=
            source->real() * s + source->imag() * c;
      ++source;
      ++sqrtLMIter;
    }
  }
}

// gridding/wstackinggridder.cpp:323-336 and 200-239. Layers are assigned to
// FFT threads round-robin instead of through a shared stack, which keeps the
// per-thread accumulation order deterministic.
template <typename T>
void WStackingRef<T>::FinishInversionPass() {
  const size_t layerOffset = LayerRangeStart(pass_);
  const size_t nLayers = layers_.size();
  const size_t nThreads = std::max<size_t>(1, std::min(nFFTThreads_, nLayers));
  auto worker = [&](size_t t) {
    Fft2D<T> fft(width_, height_, +1);
    std::vector<std::complex<T>> buffer(width_ * height_);
    for (size_t layer = t; layer < nLayers; layer += nThreads) {
      std::copy(layers_[layer].begin(), layers_[layer].end(), buffer.begin());
      fft.Transform(buffer.data());
      if (isComplex_)
        projectOnImageAndCorrect<true>(buffer.data(),
                                       LayerToW(layer + layerOffset), t);
      else
        projectOnImageAndCorrect<false>(buffer.data(),
                                        LayerToW(layer + layerOffset), t);
    }
  };
  std::vector<std::thread> threads;
  for (size_t t = 0; t != nThreads; ++t) threads.emplace_back(worker, t);
  for (auto& th : threads) th.join();
}

template <typename T>
std::vector<double> WStackingRef<T>::KernelCorrection(size_t axis) const {
  // REDFT01 of n = axis*OS/2 points, input h[j] = k1d[K*OS/2 + j] for
  // j = 0..K*OS/2 and zero beyond: D[k] = h0 + 2 sum_{j>=1} h_j cos(pi j (k+1/2)/n)
  const size_t n = axis * overSampling_ / 2;
  const size_t half = kernelSize_ * overSampling_ / 2;
  std::vector<double> out(axis / 2 + 1);
  for (size_t k = 0; k != out.size(); ++k) {
    long double acc = k1d_[half];
    for (size_t j = 1; j <= half && j < n; ++j)
      acc += 2.0L * (long double)k1d_[half + j] *
             std::cos((long double)kPi * (long double)j * ((long double)k + 0.5L) /
                      (long double)n);
    out[k] = double(acc);
  }
  return out;
}

// gridding/wstackinggridder.cpp:717-827. T=float rounds the DCT output to
// float before use, as the fftwf variant does.
template <typename T>
template <bool Inverse>
void WStackingRef<T>::correctImageForKernel(T* image) const {
  const std::vector<double> dx = KernelCorrection(width_);
  const std::vector<double> dy =
      (width_ == height_) ? dx : KernelCorrection(height_);
  const double normFactor = 1.0 / (overSampling_ * overSampling_);
  for (size_t y = 0; y != height_; ++y) {
    for (size_t x = 0; x != width_; ++x) {
      const T xVal = T((x >= width_ / 2) ? dx[x - width_ / 2] : dx[width_ / 2 - x]);
      const T yVal =
          T((y >= height_ / 2) ? dy[y - height_ / 2] : dy[height_ / 2 - y]);
      if (Inverse)
        *image *= xVal * yVal * normFactor;
      else
        *image /= xVal * yVal * normFactor;
      ++image;
    }
  }
}

// gridding/wstackinggridder.cpp:698-715
template <typename T>
void WStackingRef<T>::finalizeImage(double mult,
                                    std::vector<std::vector<T>>& data) {
  for (size_t i = 1; i < nFFTThreads_; ++i) {
    T* primary = data[0].data();
    for (T value : data[i]) {
      *primary += value;
      ++primary;
    }
  }
  for (T& value : data[0]) value *= mult;
  if (gridMode_ != KernelMode::NearestNeighbour)
    correctImageForKernel<false>(data[0].data());
}

// gridding/wstackinggridder.cpp:691-696
template <typename T>
void WStackingRef<T>::FinalizeImage(double mult) {
  layers_.clear();
  finalizeImage(mult, image_);
  if (isComplex_) finalizeImage(mult, imageImag_);
}

// gridding/wstackinggridder.cpp:829-849
template <typename T>
void WStackingRef<T>::initializePrediction(const float* in,
                                           std::vector<std::vector<T>>& data) {
  if (data[0].size() != width_ * height_) data[0].assign(width_ * height_, T(0));
  T* dataPtr = data[0].data();
  for (size_t y = 0; y != height_; ++y) {
    const double m = ((double)y - (height_ / 2)) * pixelSizeY_ + mShift_;
    for (size_t x = 0; x != width_; ++x) {
      const double l = ((width_ / 2) - (double)x) * pixelSizeX_ + lShift_;
      if (std::isfinite(*dataPtr) && l * l + m * m < 1.0)
        *dataPtr = *in;
      else
        *dataPtr = 0.0;
      ++dataPtr;
      ++in;
    }
  }
  if (gridMode_ != KernelMode::NearestNeighbour)
    correctImageForKernel<false>(data[0].data());
}

template <typename T>
void WStackingRef<T>::InitializePrediction(const float* real,
                                           const float* imag) {
  initializePrediction(real, image_);
  if (imag) initializePrediction(imag, imageImag_);
}

// gridding/wstackinggridder.cpp:934-969
template <typename T>
template <bool IsComplexImpl>
void WStackingRef<T>::copyImageToLayerAndInverseCorrect(std::complex<T>* dest,
                                                        double w) {
  const T* dataReal = image_[0].data();
  const T* dataImaginary = IsComplexImpl ? imageImag_[0].data() : nullptr;
  const double twoPiW = 2.0 * kPi * w;
  const T* sqrtLMIter = sqrtLM_.data();
  for (size_t y = 0; y != height_; ++y) {
    size_t yDest = y + height_ / 2;
    if (yDest >= height_) yDest -= height_;
    for (size_t x = 0; x != width_; ++x) {
      size_t xDest = (width_ - x) + width_ / 2;
      if (xDest >= width_) xDest -= width_;
      const double rad = twoPiW * *sqrtLMIter, s = std::sin(rad),
                   c = std::cos(rad);
      const T realVal = dataReal[xDest + yDest * width_];
      if (IsComplexImpl) {
        const T imagVal = -dataImaginary[xDest + yDest * width_];
        *dest = std::complex<T>(realVal * c + imagVal * s,
                                imagVal * c - realVal * s);
      } else {
        *dest = std::complex<T>(realVal * c, -realVal * s);
      }
      ++dest;
      ++sqrtLMIter;
    }
  }
}

// gridding/wstackinggridder.cpp:139-157 and 282-321
template <typename T>
void WStackingRef<T>::StartPredictionPass(size_t pass) {
  initializeSqrtLMLookupTableForSampling();
  pass_ = pass;
  const size_t layerOffset = LayerRangeStart(pass);
  const size_t nLayers = LayerRangeStart(pass + 1) - layerOffset;
  layers_.resize(nLayers);
  for (auto& l : layers_) l.resize(width_ * height_);
  const size_t nThreads = std::max<size_t>(1, std::min(nFFTThreads_, nLayers));
  auto worker = [&](size_t t) {
    Fft2D<T> fft(width_, height_, -1);
    for (size_t layer = t; layer < nLayers; layer += nThreads) {
      if (isComplex_)
        copyImageToLayerAndInverseCorrect<true>(layers_[layer].data(),
                                                LayerToW(layer + layerOffset));
      else
        copyImageToLayerAndInverseCorrect<false>(layers_[layer].data(),
                                                 LayerToW(layer + layerOffset));
      fft.Transform(layers_[layer].data());
    }
  };
  std::vector<std::thread> threads;
  for (size_t t = 0; t != nThreads; ++t) threads.emplace_back(worker, t);
  for (auto& th : threads) th.join();
}

// gridding/wstackinggridder.cpp:599-689
template <typename T>
void WStackingRef<T>::SampleDataSample(std::complex<float>& value, double u,
                                       double v, double w) const {
  const float nan = std::numeric_limits<float>::quiet_NaN();
  const size_t layerOffset = LayerRangeStart(pass_),
               layerRangeEnd = LayerRangeStart(pass_ + 1);
  const bool isConjugated = (w < 0.0 && !isComplex_);
  if (isConjugated) {
    u = -u;
    v = -v;
    w = -w;
  }
  const size_t wLayer = WToLayer(w);
  if (!(wLayer >= layerOffset && wLayer < layerRangeEnd)) {
    value = std::complex<float>(nan, nan);
    return;
  }
  const std::complex<T>* uvData = layers_[wLayer - layerOffset].data();
  const int W = int(width_), H = int(height_);
  std::complex<float> sample;
  if (gridMode_ == KernelMode::NearestNeighbour) {
    int x = int(std::round(u * pixelSizeX_ * width_)),
        y = int(std::round(v * pixelSizeY_ * height_));
    if (x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2) {
      if (x < 0) x += W;
      if (y < 0) y += H;
      sample = std::complex<float>(float(uvData[x + y * width_].real()),
                                   float(uvData[x + y * width_].imag()));
    } else {
      sample = std::complex<float>(nan, nan);
    }
  } else {
    sample = 0.0;
    const double xExact = u * pixelSizeX_ * width_,
                 yExact = v * pixelSizeY_ * height_;
    int x = int(std::round(xExact)), y = int(std::round(yExact));
    const int os = int(overSampling_);
    int xKernelIndex = int(std::round((xExact - double(x)) * overSampling_));
    int yKernelIndex = int(std::round((yExact - double(y)) * overSampling_));
    xKernelIndex = (xKernelIndex + (os * 3) / 2) % os;
    yKernelIndex = (yKernelIndex + (os * 3) / 2) % os;
    const std::vector<T>& xKernel = kernels_[xKernelIndex];
    const std::vector<T>& yKernel = kernels_[yKernelIndex];
    const int mid = int(kernelSize_) / 2;
    if (x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2) {
      if (x < 0) x += W;
      if (y < 0) y += H;
      for (size_t j = 0; j != kernelSize_; ++j) {
        const T yKernelValue = yKernel[j];
        const size_t cy = ((y + j + height_ - mid) % height_) * width_;
        for (size_t i = 0; i != kernelSize_; ++i) {
          const T kernelValue = xKernel[i] * yKernelValue;
          const size_t cx = (x + i + width_ - mid) % width_;
          const std::complex<T>& g = uvData[cx + cy];
          sample += std::complex<float>(float(g.real() * kernelValue),
                                        float(g.imag() * kernelValue));
        }
      }
    } else {
      sample = std::complex<float>(nan, nan);
    }
  }
  value = isConjugated ? sample : std::conj(sample);
}

template class WStackingRef<float>;
template class WStackingRef<double>;

}  // namespace oracle
