// ORACLE / TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "msref.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <thread>

namespace oracle {

namespace {
constexpr double kPi = 3.14159265358979323846;

// Bounded multi-producer/multi-consumer queue with the write/read/write_end
// protocol of aocommon::Lane as used at gridding/wsmsgridder.cpp:144-253.
// Items are moved in batches, as the lane_write_buffer/lane_read_buffer
// helpers of system/buffered_lane.h:8-97 do.
template <typename Item>
class BatchLane {
 public:
  explicit BatchLane(size_t capacityBatches) : capacity_(capacityBatches) {}
  void Write(std::vector<Item>&& batch) {
    std::unique_lock<std::mutex> lock(mutex_);
    notFull_.wait(lock, [&] { return queue_.size() < capacity_; });
    queue_.push_back(std::move(batch));
    notEmpty_.notify_one();
  }
  bool Read(std::vector<Item>& batch) {
    std::unique_lock<std::mutex> lock(mutex_);
    notEmpty_.wait(lock, [&] { return !queue_.empty() || ended_; });
    if (queue_.empty()) return false;
    batch = std::move(queue_.front());
    queue_.pop_front();
    notFull_.notify_one();
    return true;
  }
  void WriteEnd() {
    std::lock_guard<std::mutex> lock(mutex_);
    ended_ = true;
    notEmpty_.notify_all();
  }

 private:
  size_t capacity_;
  std::deque<std::vector<Item>> queue_;
  bool ended_ = false;
  std::mutex mutex_;
  std::condition_variable notEmpty_, notFull_;
};

template <typename Item>
class BufferedWriter {
 public:
  BufferedWriter(BatchLane<Item>* lane, size_t size) : lane_(lane), size_(size) {
    buffer_.reserve(size);
  }
  void Write(const Item& item) {
    buffer_.push_back(item);
    if (buffer_.size() == size_) flush();
  }
  void End() {
    flush();
    lane_->WriteEnd();
  }

 private:
  void flush() {
    if (!buffer_.empty()) {
      lane_->Write(std::move(buffer_));
      buffer_ = std::vector<Item>();
      buffer_.reserve(size_);
    }
  }
  BatchLane<Item>* lane_;
  size_t size_;
  std::vector<Item> buffer_;
};

struct Sample {
  double u, v, w;
  std::complex<float> value;
};

// gridding/msgridderbase.cpp:692-704
void rotate(const MsView& ms, double shiftFactor, std::complex<float>* data) {
  for (size_t ch = 0; ch != ms.nchan; ++ch) {
    const double wShiftRad = shiftFactor / ms.wavelengths[ch];
    const std::complex<float> phasor(std::cos(wShiftRad), std::sin(wShiftRad));
    data[ch] *= phasor;
  }
}

double smallestWavelength(const MsView& ms) {
  return *std::min_element(ms.wavelengths, ms.wavelengths + ms.nchan);
}
double longestWavelength(const MsView& ms) {
  return *std::max_element(ms.wavelengths, ms.wavelengths + ms.nchan);
}

}  // namespace

// aocommon::Image::Trim / Untrim keep the centre of the image.
std::vector<float> trimImage(const std::vector<float>& in, size_t w, size_t h,
                             size_t tw, size_t th) {
  std::vector<float> out(tw * th);
  const size_t sx = (w - tw) / 2, sy = (h - th) / 2;
  for (size_t y = 0; y != th; ++y)
    for (size_t x = 0; x != tw; ++x) out[x + y * tw] = in[(x + sx) + (y + sy) * w];
  return out;
}
std::vector<float> untrimImage(const float* in, size_t tw, size_t th, size_t w,
                               size_t h) {
  std::vector<float> out(w * h, 0.0f);
  const size_t sx = (w - tw) / 2, sy = (h - th) / 2;
  for (size_t y = 0; y != th; ++y)
    for (size_t x = 0; x != tw; ++x) out[(x + sx) + (y + sy) * w] = in[x + y * tw];
  return out;
}

// ---------------------------------------------------------------- ImageWeights
ImageWeightsRef::ImageWeightsRef(WeightingMode mode, double briggsRobustness,
                                 size_t width, size_t height,
                                 double pixelScaleX, double pixelScaleY,
                                 bool weightsAsTaper, double superWeight)
    : mode_(mode),
      robust_(briggsRobustness),
      width_(size_t(std::round(double(width) / superWeight))),
      height_(size_t(std::round(double(height) / superWeight))),
      scaleX_(pixelScaleX),
      scaleY_(pixelScaleY),
      weightsAsTaper_(weightsAsTaper) {
  if (width_ % 2 != 0) ++width_;
  if (height_ % 2 != 0) ++height_;
  grid_.assign(width_ * height_ / 2, 0.0);
}

// structures/imageweights.h:73-80
void ImageWeightsRef::uvToXY(double u, double v, int& x, int& y) const {
  if (v < 0.0) {
    u = -u;
    v = -v;
  }
  x = int(std::floor(u * width_ * scaleX_ + width_ / 2));
  y = int(std::floor(v * height_ * scaleY_));
}

// structures/imageweights.h:92-94
bool ImageWeightsRef::isWithinLimits(int x, int y) const {
  return x >= 0 && x < int(width_) && y < int(height_ / 2);
}

double ImageWeightsRef::GetWeight(double u, double v) const {
  int x, y;
  uvToXY(u, v, x, y);
  if (isWithinLimits(x, y)) return grid_[size_t(x) + size_t(y) * width_];
  return 0.0;
}

// structures/imageweights.cpp:75-111 and imageweights.h:36-45
void ImageWeightsRef::Grid(size_t nrow, size_t nchan, size_t npol,
                           const double* uvw_m, const double* wavelengths,
                           const float* weights) {
  std::vector<float> wbuf(nchan * npol);
  for (size_t r = 0; r != nrow; ++r) {
    double uInM = uvw_m[r * 3 + 0], vInM = uvw_m[r * 3 + 1];
    for (size_t i = 0; i != nchan * npol; ++i)
      wbuf[i] = weights ? weights[r * nchan * npol + i] : 1.0f;
    if (weightsAsTaper_)
      for (float& w : wbuf)
        if (w != 0.0) w = 1.0;
    if (vInM < 0.0) {
      uInM = -uInM;
      vInM = -vInM;
    }
    const float* wIter = wbuf.data();
    for (size_t ch = 0; ch != nchan; ++ch) {
      const double u = uInM / wavelengths[ch];
      const double v = vInM / wavelengths[ch];
      for (size_t p = 0; p != npol; ++p) {
        int x, y;
        uvToXY(u, v, x, y);
        if (isWithinLimits(x, y)) {
          grid_[size_t(x) + size_t(y) * width_] += *wIter;
          totalSum_ += *wIter;
        }
        ++wIter;
      }
    }
  }
}

// structures/imageweights.cpp:113-144
void ImageWeightsRef::FinishGridding() {
  switch (mode_) {
    case WeightingMode::Briggs: {
      double avgW = 0.0;
      for (double val : grid_) avgW += val * val;
      avgW /= totalSum_;
      const double numeratorSqrt = 5.0 * std::exp((-M_LN10) * robust_);
      const double sSq = numeratorSqrt * numeratorSqrt / avgW;
      for (double& val : grid_) val = 1.0 / (1.0 + val * sSq);
    } break;
    case WeightingMode::Uniform:
      for (double& val : grid_) val = (val != 0.0) ? 1.0 / val : 0.0;
      break;
    case WeightingMode::Natural:
      for (double& val : grid_)
        if (val != 0.0) val = 1.0;
      break;
  }
}

// ------------------------------------------------------------ host metadata
void CalculateWLimits(const GridderSettings& s, const MsView& ms,
                      const ImageWeightsRef& iw, double& minWOut,
                      double& maxWOut, double& maxBaselineUVW) {
  double maxW = 0.0, minW = 1e100;
  maxBaselineUVW = 0.0;
  const double smallestWl = smallestWavelength(ms);
  const double longestWl = longestWavelength(ms);
  const double halfWidth = 0.5 * s.width, halfHeight = 0.5 * s.height;
  for (size_t r = 0; r != ms.nrow; ++r) {
    const double uInM = ms.uvw_m[r * 3], vInM = ms.uvw_m[r * 3 + 1],
                 wInM = ms.uvw_m[r * 3 + 2];
    const double wHi = std::fabs(wInM / smallestWl);
    const double wLo = std::fabs(wInM / longestWl);
    const double baselineInM =
        std::sqrt(uInM * uInM + vInM * vInM + wInM * wInM);
    if (wHi > maxW || wLo < minW || baselineInM / smallestWl > maxBaselineUVW) {
      for (size_t ch = 0; ch != ms.nchan; ++ch) {
        const float weight = ms.weights ? ms.weights[r * ms.nchan + ch] : 1.0f;
        if (weight == 0.0) continue;
        const double wl = ms.wavelengths[ch];
        const double wInL = wInM / wl, uInL = uInM / wl, vInL = vInM / wl;
        const double x = uInL * s.pixelSizeX * s.width,
                     y = vInL * s.pixelSizeY * s.height;
        const double imagingWeight = iw.GetWeight(uInL, vInL);
        if (imagingWeight != 0.0 && std::floor(x) > -halfWidth &&
            std::ceil(x) < halfWidth && std::floor(y) > -halfHeight &&
            std::ceil(y) < halfHeight) {
          maxW = std::max(maxW, std::fabs(wInL));
          minW = std::min(minW, std::fabs(wInL));
          maxBaselineUVW = std::max(maxBaselineUVW, baselineInM / wl);
        }
      }
    }
  }
  if (minW == 1e100) {
    minW = 0.0;
    maxW = 0.0;
  }
  minWOut = minW;
  maxWOut = maxW;
}

std::vector<double> Resample(const double* in, size_t inW, size_t inH, size_t outW,
                             size_t outH) {
  // forward transform of the real input (as complex: the half spectrum is
  // columns 0 .. inW / 2 of it)
  std::vector<std::complex<double>> x(inW * inH);
  for (size_t i = 0; i != inW * inH; ++i) x[i] = in[i];
  Fft2D<double> fwd(inW, inH, -1);
  fwd.Transform(x.data());
  const size_t minW = std::min(inW / 2 + 1, outW / 2 + 1);
  const size_t minH = std::min(inH, outH);
  // the copied half spectrum: rows 0 .. minH/2 - 1 to the same rows, the last
  // minH/2 input rows to the last minH/2 output rows
  std::vector<std::complex<double>> y(outH * minW, 0.0);
  for (size_t r = 0; r != minH / 2; ++r) {
    const size_t oldY = inH - minH / 2 + r, newY = outH - minH / 2 + r;
    for (size_t c = 0; c != minW; ++c) y[newY * minW + c] = x[oldY * inW + c];
  }
  for (size_t r = 0; r != minH / 2; ++r)
    for (size_t c = 0; c != minW; ++c) y[r * minW + c] = x[r * inW + c];
  // complex-to-real transform of the out_w x out_h Hermitian half spectrum:
  // inverse transform of each column along y, then along x the real part of
  // sum_k c_k z_k exp(2 pi i x k / out_w), c_0 = 1, c_k = 2 below the
  // output's Nyquist column, 1 at it
  std::vector<std::complex<double>> col(outH);
  Fft1D<double> colInv(outH, +1);
  std::vector<std::complex<double>> scratch(outH);
  for (size_t c = 0; c != minW; ++c) {
    for (size_t r = 0; r != outH; ++r) col[r] = y[r * minW + c];
    colInv.Transform(col.data(), scratch.data());
    for (size_t r = 0; r != outH; ++r) y[r * minW + c] = col[r];
  }
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<std::complex<double>> tw(outW);
  for (size_t k = 0; k != outW; ++k) {
    const long double a = 2.0L * pi * (long double)k / (long double)outW;
    tw[k] = std::complex<double>(double(std::cos(a)), double(std::sin(a)));
  }
  std::vector<double> out(outW * outH);
  const double factor = 1.0 / (double(inW) * double(inH));
  for (size_t r = 0; r != outH; ++r)
    for (size_t xo = 0; xo != outW; ++xo) {
      double acc = 0.0;
      for (size_t c = 0; c != minW; ++c) {
        const double ck = (c == 0 || 2 * c == outW) ? 1.0 : 2.0;
        acc += ck * (y[r * minW + c] * tw[(xo * c) % outW]).real();
      }
      out[r * outW + xo] = acc * factor;
    }
  return out;
}

size_t SuggestedWGridSize(const GridderSettings& s, double minW, double maxW) {
  const size_t tw = s.trimWidth ? s.trimWidth : s.width;
  const size_t th = s.trimHeight ? s.trimHeight : s.height;
  const double maxL = tw * s.pixelSizeX * 0.5 + std::fabs(s.lShift),
               maxM = th * s.pixelSizeY * 0.5 + std::fabs(s.mShift),
               lmSq = maxL * maxL + maxM * maxM;
  const double cMinW = s.isComplex ? -maxW : minW;
  double radiansForAllLayers;
  if (lmSq < 1.0)
    radiansForAllLayers = 2 * kPi * (maxW - cMinW) * (1.0 - std::sqrt(1.0 - lmSq));
  else
    radiansForAllLayers = 2 * kPi * (maxW - cMinW);
  size_t suggested = size_t(std::ceil(radiansForAllLayers * s.nWLayersFactor));
  if (suggested == 0) suggested = 1;
  if (suggested < s.nCpus) {
    const double memoryRequired =
        double(s.nCpus) * double(sizeof(float)) * double(s.width * s.height);
    if (4.0 * memoryRequired < s.memory) suggested = s.nCpus;
  }
  return suggested;
}

void CalculateFFTSize(size_t size, double pixelScale, double beamSize,
                      size_t& minimalSize, size_t& optimalSize) {
  auto lowPrimes = [](size_t number) {
    while (number > 7) {
      if (number % 2 == 0)
        number /= 2;
      else if (number % 3 == 0)
        number /= 3;
      else if (number % 5 == 0)
        number /= 5;
      else if (number % 7 == 0)
        number /= 7;
      else
        return false;
    }
    return true;
  };
  const double totalSize = double(size) * pixelScale;
  minimalSize = size_t(std::ceil(totalSize * 2 / beamSize));
  optimalSize = minimalSize;
  if (optimalSize % 4 != 0) optimalSize += 4 - (optimalSize % 4);
  while (!lowPrimes(optimalSize) && optimalSize < size) optimalSize += 4;
  if (optimalSize > size) optimalSize = size;
}

namespace {
struct Meta {
  double minW, maxW;
  size_t nW;
  // inversion size and pixel scales (the image's unless small inversion
  // shrinks them)
  size_t width, height;
  double dl, dm;
};

// msgridderbase.cpp:370-446
Meta overallMetaData(const GridderSettings& s, const MsView& ms,
                     const ImageWeightsRef& iw) {
  double minW, maxW, maxBaseline;
  CalculateWLimits(s, ms, iw, minW, maxW, maxBaseline);
  if (minW > maxW) minW = maxW;
  const double beamSize = 1.0 / maxBaseline;
  if (s.wLimit != 0.0) {
    maxW *= (1.0 - s.wLimit / 100.0);
    if (maxW < minW) maxW = minW;
  }
  Meta m{minW, maxW, 0, s.width, s.height, s.pixelSizeX, s.pixelSizeY};
  if (s.smallInversion) {
    size_t optW, optH, minWd, minHt;
    CalculateFFTSize(m.width, m.dl, beamSize, minWd, optW);
    CalculateFFTSize(m.height, m.dm, beamSize, minHt, optH);
    if (optW < m.width || optH < m.height) {
      const size_t nw = std::max(std::min(optW, m.width), size_t(32));
      const size_t nh = std::max(std::min(optH, m.height), size_t(32));
      m.dl = (double(m.width) * m.dl) / double(nw);
      m.dm = (double(m.height) * m.dm) / double(nh);
      m.width = nw;
      m.height = nh;
    }
  }
  GridderSettings sm = s;
  sm.width = m.width;
  sm.height = m.height;
  m.nW = s.nWLayers != 0 ? s.nWLayers : SuggestedWGridSize(sm, minW, maxW);
  return m;
}
}  // namespace

// gridding/wsmsgridder.cpp:367-486, with gridMeasurementSet (:130-222),
// workThreadPerSample (:242-253) and readAndWeightVisibilities<1>
// (msgridderbase.cpp:515-660).
template <typename T>
InvertResult InvertMs(const GridderSettings& s, const MsView& ms,
                      const ImageWeightsRef& iw) {
  const Meta meta = overallMetaData(s, ms, iw);
  WStackingRef<T> gridder(meta.width, meta.height, meta.dl, meta.dm,
                          s.nCpus, s.kernelSize, s.overSampling);
  gridder.SetGridMode(s.gridMode);
  if (s.lShift != 0.0 || s.mShift != 0.0)
    gridder.SetDenormalPhaseCentre(s.lShift, s.mShift);
  gridder.SetIsComplex(s.isComplex);
  gridder.PrepareWLayers(meta.nW, s.memory * (6.0 / 10.0), meta.minW,
                         meta.maxW);

  InvertResult res;
  res.minW = meta.minW;
  res.maxW = meta.maxW;
  res.actualWGridSize = meta.nW;
  res.nPasses = gridder.NPasses();
  const size_t nCpus = std::max<size_t>(1, s.nCpus);
  const double smallestWl = smallestWavelength(ms), longestWl = longestWavelength(ms);
  const bool hasShift = s.lShift != 0.0 || s.mShift != 0.0;
  std::vector<std::complex<float>> rowData(ms.nchan);
  std::vector<float> weightBuffer(ms.nchan), imageWeights(ms.nchan);
  std::vector<char> isSelected(ms.nchan);

  for (size_t pass = 0; pass != gridder.NPasses(); ++pass) {
    gridder.StartInversionPass(pass);
    // startInversionWorkThreads, wsmsgridder.cpp:223-234
    const size_t laneBuffer = std::max<size_t>(nCpus * 2, 1024);
    const size_t capacity = ms.nchan * laneBuffer;
    size_t bufferSize = std::max<size_t>(8u, capacity / 8);
    bufferSize = std::min<size_t>(128, std::min(bufferSize, capacity));
    std::vector<std::unique_ptr<BatchLane<Sample>>> lanes;
    std::vector<std::thread> workers;
    for (size_t i = 0; i != nCpus; ++i)
      lanes.emplace_back(new BatchLane<Sample>(
          std::max<size_t>(1, capacity / bufferSize)));
    for (size_t i = 0; i != nCpus; ++i)
      workers.emplace_back([&, i] {
        std::vector<Sample> batch;
        while (lanes[i]->Read(batch))
          for (const Sample& smp : batch)
            gridder.AddDataSample(smp.value, smp.u, smp.v, smp.w);
      });
    std::vector<BufferedWriter<Sample>> writers;
    for (size_t i = 0; i != nCpus; ++i)
      writers.emplace_back(lanes[i].get(), bufferSize);

    for (size_t r = 0; r != ms.nrow; ++r) {
      const double uInM = ms.uvw_m[r * 3], vInM = ms.uvw_m[r * 3 + 1],
                   wInM = ms.uvw_m[r * 3 + 2];
      const double w1 = wInM / longestWl, w2 = wInM / smallestWl;
      if (!gridder.IsInLayerRange(w1, w2)) continue;
      for (size_t ch = 0; ch != ms.nchan; ++ch)
        isSelected[ch] = gridder.IsInLayerRange(wInM / ms.wavelengths[ch]);
      // readAndWeightVisibilities<1>
      for (size_t ch = 0; ch != ms.nchan; ++ch) {
        rowData[ch] = s.psfMode ? std::complex<float>(1.0f, 0.0f)
                                : ms.data[r * ms.nchan + ch];
        weightBuffer[ch] = ms.weights ? ms.weights[r * ms.nchan + ch] : 1.0f;
        if (!isSelected[ch]) weightBuffer[ch] = 0.0;
        imageWeights[ch] = float(iw.GetWeight(uInM / ms.wavelengths[ch],
                                              vInM / ms.wavelengths[ch]));
      }
      for (size_t ch = 0; ch != ms.nchan; ++ch) {
        const double cumWeight = weightBuffer[ch] * imageWeights[ch];
        if (cumWeight != 0.0) {
          res.visibilityWeightSum += weightBuffer[ch];
          res.maxGriddedWeight = std::max(cumWeight, res.maxGriddedWeight);
          ++res.griddedVisibilityCount;
          res.totalWeight += cumWeight;
        }
        rowData[ch] *= float(cumWeight);
      }
      if (hasShift) {
        const double shiftFactor = -2.0 * kPi * (uInM * s.lShift + vInM * s.mShift);
        rotate(ms, shiftFactor, rowData.data());
      }
      for (size_t ch = 0; ch != ms.nchan; ++ch) {
        const double wl = ms.wavelengths[ch];
        Sample smp{uInM / wl, vInM / wl, wInM / wl, rowData[ch]};
        const size_t cpu = gridder.WToLayer(smp.w) % nCpus;
        writers[cpu].Write(smp);
      }
    }
    for (auto& wr : writers) wr.End();
    for (auto& th : workers) th.join();
    gridder.FinishInversionPass();
  }
  gridder.FinalizeImage(1.0 / res.totalWeight);
  const size_t n = s.width * s.height;
  std::vector<float> real(n), imag;
  // wsmsgridder.cpp:452-474: the inversion-size image interpolated to the
  // image size
  auto toImage = [&](const T* img, std::vector<float>& out) {
    if (meta.width == s.width && meta.height == s.height) {
      for (size_t i = 0; i != n; ++i) out[i] = float(img[i]);
      return;
    }
    std::vector<double> in(meta.width * meta.height);
    for (size_t i = 0; i != in.size(); ++i) in[i] = double(float(img[i]));
    const std::vector<double> r = Resample(in.data(), meta.width, meta.height, s.width, s.height);
    for (size_t i = 0; i != n; ++i) out[i] = float(r[i]);
  };
  toImage(gridder.RealImage(), real);
  if (s.isComplex) {
    imag.resize(n);
    toImage(gridder.ImaginaryImage(), imag);
  }
  const size_t tw = s.trimWidth ? s.trimWidth : s.width;
  const size_t th = s.trimHeight ? s.trimHeight : s.height;
  if (tw != s.width || th != s.height) {
    real = trimImage(real, s.width, s.height, tw, th);
    if (s.isComplex) imag = trimImage(imag, s.width, s.height, tw, th);
  }
  res.real = std::move(real);
  res.imag = std::move(imag);
  return res;
}

// gridding/wsmsgridder.cpp:488-586 with predictMeasurementSet (:255-315),
// predictCalcThread (:317-336) and predictWriteThread (:338-365).
template <typename T>
void PredictMs(const GridderSettings& s, const MsView& ms,
               const ImageWeightsRef& iw, const float* real, const float* imag,
               std::complex<float>* out, double* minWOut, double* maxWOut,
               size_t* nPassesOut) {
  if (s.isComplex && !imag)
    throw std::runtime_error("Missing imaginary in complex prediction");
  if (!s.isComplex && imag)
    throw std::runtime_error("Imaginary specified in non-complex prediction");
  const Meta meta = overallMetaData(s, ms, iw);
  WStackingRef<T> gridder(meta.width, meta.height, meta.dl, meta.dm,
                          s.nCpus, s.kernelSize, s.overSampling);
  gridder.SetGridMode(s.gridMode);
  const bool hasShift = s.lShift != 0.0 || s.mShift != 0.0;
  if (hasShift) gridder.SetDenormalPhaseCentre(s.lShift, s.mShift);
  gridder.SetIsComplex(s.isComplex);
  gridder.PrepareWLayers(meta.nW, s.memory * (6.0 / 10.0), meta.minW,
                         meta.maxW);
  if (minWOut) *minWOut = meta.minW;
  if (maxWOut) *maxWOut = meta.maxW;
  if (nPassesOut) *nPassesOut = gridder.NPasses();
  const size_t tw = s.trimWidth ? s.trimWidth : s.width;
  const size_t th = s.trimHeight ? s.trimHeight : s.height;
  std::vector<float> re, im;
  if (tw != s.width || th != s.height) {
    re = untrimImage(real, tw, th, s.width, s.height);
    if (imag) im = untrimImage(imag, tw, th, s.width, s.height);
  } else {
    re.assign(real, real + s.width * s.height);
    if (imag) im.assign(imag, imag + s.width * s.height);
  }
  // wsmsgridder.cpp:521-551: the image-size model decimated to the
  // inversion size
  if (meta.width != s.width || meta.height != s.height) {
    auto down = [&](std::vector<float>& img) {
      std::vector<double> in(img.begin(), img.end());
      const std::vector<double> r =
          Resample(in.data(), s.width, s.height, meta.width, meta.height);
      img.assign(r.size(), 0.0f);
      for (size_t i = 0; i != r.size(); ++i) img[i] = float(r[i]);
    };
    down(re);
    if (imag) down(im);
  }
  gridder.InitializePrediction(re.data(), imag ? im.data() : nullptr);
  const size_t nCpus = std::max<size_t>(1, s.nCpus);
  const size_t laneBuffer = std::max<size_t>(nCpus * 2, 1024);

  struct Item {
    size_t row;
    std::vector<std::complex<float>> data;
  };
  for (size_t pass = 0; pass != gridder.NPasses(); ++pass) {
    gridder.StartPredictionPass(pass);
    BatchLane<Item> calcLane(std::max<size_t>(2, (laneBuffer + nCpus) / 64));
    BatchLane<Item> writeLane(std::max<size_t>(2, laneBuffer / 64));
    std::thread writer([&] {
      auto cmp = [](const Item& a, const Item& b) { return a.row > b.row; };
      std::priority_queue<Item, std::vector<Item>, decltype(cmp)> queue(cmp);
      size_t next = 0;
      std::vector<Item> batch;
      while (writeLane.Read(batch)) {
        for (Item& it : batch) queue.emplace(std::move(it));
        while (!queue.empty() && queue.top().row == next) {
          const Item& it = queue.top();
          for (size_t ch = 0; ch != ms.nchan; ++ch) {
            const std::complex<float> v = it.data[ch];
            if (std::isfinite(v.real()) && std::isfinite(v.imag()))
              out[it.row * ms.nchan + ch] = v;
          }
          queue.pop();
          ++next;
        }
      }
    });
    std::vector<std::thread> calc;
    std::mutex writeMutex;
    for (size_t t = 0; t != nCpus; ++t)
      calc.emplace_back([&] {
        std::vector<Item> batch;
        while (calcLane.Read(batch)) {
          for (Item& it : batch) {
            const double uInM = ms.uvw_m[it.row * 3],
                         vInM = ms.uvw_m[it.row * 3 + 1],
                         wInM = ms.uvw_m[it.row * 3 + 2];
            for (size_t ch = 0; ch != ms.nchan; ++ch) {
              const double wl = ms.wavelengths[ch];
              gridder.SampleDataSample(it.data[ch], uInM / wl, vInM / wl,
                                       wInM / wl);
            }
            if (hasShift) {
              const double shiftFactor =
                  2.0 * kPi * (uInM * s.lShift + vInM * s.mShift);
              rotate(ms, shiftFactor, it.data.data());
            }
          }
          writeLane.Write(std::move(batch));
          batch = std::vector<Item>();
        }
      });
    std::vector<Item> batch;
    for (size_t r = 0; r != ms.nrow; ++r) {
      batch.push_back(Item{r, std::vector<std::complex<float>>(ms.nchan)});
      if (batch.size() == 64) {
        calcLane.Write(std::move(batch));
        batch = std::vector<Item>();
      }
    }
    if (!batch.empty()) calcLane.Write(std::move(batch));
    calcLane.WriteEnd();
    for (auto& t : calc) t.join();
    writeLane.WriteEnd();
    writer.join();
  }
  (void)iw;
}

template InvertResult InvertMs<float>(const GridderSettings&, const MsView&,
                                      const ImageWeightsRef&);
template InvertResult InvertMs<double>(const GridderSettings&, const MsView&,
                                       const ImageWeightsRef&);
template void PredictMs<float>(const GridderSettings&, const MsView&,
                               const ImageWeightsRef&, const float*,
                               const float*, std::complex<float>*, double*,
                               double*, size_t*);
template void PredictMs<double>(const GridderSettings&, const MsView&,
                                const ImageWeightsRef&, const float*,
                                const float*, std::complex<float>*, double*,
                                double*, size_t*);

// ------------------------------------------------------------------ FP64 DFT
namespace {
// gridding/directmsgridder.cpp:175-196: n' in the FFT-ordered (pre-remap)
// index space; pixel (x,y) of the layer maps to image pixel (xSrc, ySrc).
void lmTables(size_t width, size_t height, double dl, double dm, double ls,
              double ms, std::vector<double>& l, std::vector<double>& m,
              std::vector<double>& n) {
  l.resize(width);
  m.resize(height);
  n.resize(width * height);
  for (size_t y = 0; y != height; ++y) {
    size_t ySrc = (height - y) + height / 2;
    if (ySrc >= height) ySrc -= height;
    m[y] = ((double)ySrc - (height / 2)) * dm + ms;
  }
  for (size_t x = 0; x != width; ++x) {
    size_t xSrc = x + width / 2;
    if (xSrc >= width) xSrc -= width;
    l[x] = ((width / 2) - (double)xSrc) * dl + ls;
  }
  for (size_t y = 0; y != height; ++y)
    for (size_t x = 0; x != width; ++x) {
      const double ll = l[x], mm = m[y];
      n[x + y * width] =
          (ll * ll + mm * mm < 1.0) ? std::sqrt(1.0 - ll * ll - mm * mm) - 1.0 : 0.0;
    }
}
}  // namespace

// gridding/directmsgridder.cpp:16-114. The reference indexes rows with
// y*height (:95), which is only correct for square images; y*width is used.
void DftInvert(size_t width, size_t height, double dl, double dm,
               double lShift, double mShift, size_t n, const double* uvw,
               const std::complex<float>* vis, double factor, double* image,
               size_t nThreads) {
  std::vector<double> l, m, nn;
  lmTables(width, height, dl, dm, lShift, mShift, l, m, nn);
  nThreads = std::max<size_t>(1, nThreads);
  std::vector<std::vector<double>> layers(nThreads,
                                          std::vector<double>(width * height, 0.0));
  auto worker = [&](size_t t) {
    std::vector<double>& layer = layers[t];
    const double minTwoPi = -2.0 * kPi;
    for (size_t k = t; k < n; k += nThreads) {
      const double u = uvw[k * 3], v = uvw[k * 3 + 1], w = uvw[k * 3 + 2];
      const std::complex<double> val(vis[k].real(), vis[k].imag());
      for (size_t y = 0; y != height; ++y)
        for (size_t x = 0; x != width; ++x) {
          const size_t index = y * width + x;
          const double angle = minTwoPi * (u * l[x] + v * m[y] + w * nn[index]);
          layer[index] += val.real() * std::cos(angle) - val.imag() * std::sin(angle);
        }
    }
  };
  std::vector<std::thread> threads;
  for (size_t t = 0; t != nThreads; ++t) threads.emplace_back(worker, t);
  for (auto& th : threads) th.join();
  for (size_t t = 1; t < nThreads; ++t)
    for (size_t i = 0; i != width * height; ++i) layers[0][i] += layers[t][i];
  for (size_t y = 0; y != height; ++y) {
    size_t ySrc = (height - y) + height / 2;
    if (ySrc >= height) ySrc -= height;
    for (size_t x = 0; x != width; ++x) {
      size_t xSrc = x + width / 2;
      if (xSrc >= width) xSrc -= width;
      image[x + y * width] = layers[0][xSrc + ySrc * width] * factor;
    }
  }
}

void DftPredict(size_t width, size_t height, double dl, double dm,
                double lShift, double mShift, size_t n, const double* uvw,
                const double* image, std::complex<double>* vis,
                size_t nThreads) {
  std::vector<double> l, m, nn;
  lmTables(width, height, dl, dm, lShift, mShift, l, m, nn);
  // image pixel (x, y) corresponds to layer index (xSrc, ySrc) of the remap
  std::vector<double> layerImage(width * height);
  for (size_t y = 0; y != height; ++y) {
    size_t ySrc = (height - y) + height / 2;
    if (ySrc >= height) ySrc -= height;
    for (size_t x = 0; x != width; ++x) {
      size_t xSrc = x + width / 2;
      if (xSrc >= width) xSrc -= width;
      layerImage[xSrc + ySrc * width] = image[x + y * width];
    }
  }
  nThreads = std::max<size_t>(1, nThreads);
  auto worker = [&](size_t t) {
    for (size_t k = t; k < n; k += nThreads) {
      const double u = uvw[k * 3], v = uvw[k * 3 + 1], w = uvw[k * 3 + 2];
      double re = 0.0, im = 0.0;
      for (size_t y = 0; y != height; ++y)
        for (size_t x = 0; x != width; ++x) {
          const size_t index = y * width + x;
          const double p = layerImage[index];
          if (p == 0.0) continue;
          const double angle = 2.0 * kPi * (u * l[x] + v * m[y] + w * nn[index]);
          re += p * std::cos(angle);
          im += p * std::sin(angle);
        }
      vis[k] = std::complex<double>(re, im);
    }
  };
  std::vector<std::thread> threads;
  for (size_t t = 0; t != nThreads; ++t) threads.emplace_back(worker, t);
  for (auto& th : threads) th.join();
}

}  // namespace oracle
