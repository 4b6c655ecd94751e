#include "msgridderbase.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace wsclean_amd {

MSGridderBase::MSGridderBase(const Settings& settings)
    : _settings(settings),
      pixelSizeX_(settings.pixelScaleX),
      pixelSizeY_(settings.pixelScaleY),
      wGridSize_(settings.nWLayers),
      wLimit_(settings.wLimit / 100.0),
      visibilityWeightingMode_(settings.visibilityWeightingMode) {}

// gridding/msgridderbase.cpp:200-286 for a single-correlation provider.
void MSGridderBase::calculateWLimits(MSData& msData) {
  MSProvider& provider = *msData.msProvider;
  const aocommon::BandData band = msData.SelectedBand();
  const size_t nchan = band.ChannelCount();
  msData.maxW = 0.0;
  msData.maxWWithFlags = 0.0;
  msData.minW = 1e100;
  msData.maxBaselineUVW = 0.0;
  msData.maxBaselineInM = 0.0;
  std::vector<float> weights(provider.NChannels());
  const double smallestWl = band.SmallestWavelength(), longestWl = band.LongestWavelength();
  double curTimestep = -1, firstTime = -1, lastTime = -1;
  size_t nTimesteps = 0;
  const double halfWidth = 0.5 * ImageWidth(), halfHeight = 0.5 * ImageHeight();
  std::unique_ptr<MSReader> reader = provider.MakeReader();
  while (reader->CurrentRowAvailable()) {
    MSProvider::MetaData meta;
    reader->ReadMeta(meta);
    if (curTimestep != meta.time) {
      curTimestep = meta.time;
      ++nTimesteps;
      if (firstTime == -1) firstTime = curTimestep;
      lastTime = curTimestep;
    }
    const double wHi = std::fabs(meta.wInM / smallestWl);
    const double wLo = std::fabs(meta.wInM / longestWl);
    const double baselineInM =
        std::sqrt(meta.uInM * meta.uInM + meta.vInM * meta.vInM + meta.wInM * meta.wInM);
    if (wHi > msData.maxW || wLo < msData.minW ||
        baselineInM / smallestWl > msData.maxBaselineUVW) {
      reader->ReadWeights(weights.data());
      for (size_t ch = 0; ch != nchan; ++ch) {
        const double wl = band.ChannelWavelength(ch);
        const double wInL = meta.wInM / wl;
        msData.maxWWithFlags = std::max(msData.maxWWithFlags, std::fabs(wInL));
        if (weights[msData.startChannel + ch] != 0.0) {
          const double uInL = meta.uInM / wl, vInL = meta.vInM / wl;
          const double x = uInL * PixelSizeX() * ImageWidth();
          const double y = vInL * PixelSizeY() * ImageHeight();
          const double imagingWeight = GetImageWeights()->GetWeight(uInL, vInL);
          if (imagingWeight != 0.0 && std::floor(x) > -halfWidth && std::ceil(x) < halfWidth &&
              std::floor(y) > -halfHeight && std::ceil(y) < halfHeight) {
            msData.maxW = std::max(msData.maxW, std::fabs(wInL));
            msData.minW = std::min(msData.minW, std::fabs(wInL));
            msData.maxBaselineUVW = std::max(msData.maxBaselineUVW, baselineInM / wl);
            msData.maxBaselineInM = std::max(msData.maxBaselineInM, baselineInM);
          }
        }
      }
    }
    reader->NextInputRow();
  }
  if (msData.minW == 1e100) {
    msData.minW = 0.0;
    msData.maxWWithFlags = 0.0;
    msData.maxW = 0.0;
  }
  msData.integrationTime =
      (lastTime == firstTime || nTimesteps < 2) ? 1 : (lastTime - firstTime) / (nTimesteps - 1);
}

// gridding/msgridderbase.cpp:292-362
void MSGridderBase::initializeMSDataVector(std::vector<MSData>& msDataVector) {
  if (MeasurementSetCount() == 0)
    throw std::runtime_error(
        "Something is wrong during inversion: no measurement sets given to inversion "
        "algorithm");
  if (!imageWeights_) throw std::runtime_error("no image weights set");
  msDataVector = std::vector<MSData>(MeasurementSetCount());
  if (!_metaDataCache) _metaDataCache.reset(new MetaDataCache());
  const bool hasCache = !_metaDataCache->msDataVector.empty();
  if (!hasCache) _metaDataCache->msDataVector.resize(MeasurementSetCount());
  startTime_ = 0.0;
  for (size_t i = 0; i != MeasurementSetCount(); ++i) {
    MSData& d = msDataVector[i];
    d.msIndex = i;
    d.msProvider = measurementSets_[i];
    // the band of the provider, all channels selected (initializeBandData,
    // msgridderbase.cpp:177-198, without a channel-range selection)
    std::vector<double> wl(d.msProvider->NChannels());
    for (size_t ch = 0; ch != wl.size(); ++ch) wl[ch] = d.msProvider->ChannelWavelength(ch);
    d.bandData = aocommon::BandData(std::move(wl));
    d.startChannel = 0;
    d.endChannel = d.bandData.ChannelCount();
    // the rows the selection matches (calculateMSLimits): all of them here
    std::unique_ptr<MSReader> reader = d.msProvider->MakeReader();
    for (d.matchingRows = 0; reader->CurrentRowAvailable(); reader->NextInputRow())
      ++d.matchingRows;
    d.rowStart = 0;
    d.rowEnd = d.matchingRows;
    startTime_ = (i == 0) ? d.msProvider->StartTime() : std::min(startTime_, d.msProvider->StartTime());
    MetaDataCache::Entry& e = _metaDataCache->msDataVector[i];
    if (hasCache) {
      d.maxW = e.maxW;
      d.maxWWithFlags = e.maxWWithFlags;
      d.minW = e.minW;
      d.maxBaselineUVW = e.maxBaselineUVW;
      d.maxBaselineInM = e.maxBaselineInM;
      d.integrationTime = e.integrationTime;
    } else {
      calculateWLimits(d);
      e = MetaDataCache::Entry{d.minW, d.maxW, d.maxWWithFlags, d.maxBaselineUVW,
                               d.maxBaselineInM, d.integrationTime};
    }
  }
  calculateOverallMetaData(msDataVector);
}

// math/calculatefftsize.h:4-19
static bool lowPrimes(size_t number) {
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
}

void CalculateFFTSize(size_t size, double pixelScale, double beamSize, size_t& minimalSize,
                      size_t& optimalSize) {
  const double totalSize = double(size) * pixelScale;
  minimalSize = size_t(std::ceil(totalSize * 2 / beamSize));
  optimalSize = minimalSize;
  if (optimalSize % 4 != 0) optimalSize += 4 - (optimalSize % 4);
  while (!lowPrimes(optimalSize) && optimalSize < size) optimalSize += 4;
  if (optimalSize > size) optimalSize = size;
}

// gridding/msgridderbase.cpp:370-446
void MSGridderBase::calculateOverallMetaData(const std::vector<MSData>& msDataVector) {
  _maxW = 0.0;
  _minW = std::numeric_limits<double>::max();
  double maxBaseline = 0.0;
  for (const MSData& d : msDataVector) {
    maxBaseline = std::max(maxBaseline, d.maxBaselineUVW);
    _maxW = std::max(_maxW, d.maxW);
    _minW = std::min(_minW, d.minW);
  }
  if (_minW > _maxW) _minW = _maxW;
  theoreticalBeamSize_ = 1.0 / maxBaseline;
  if (wLimit_ != 0.0) {
    _maxW *= (1.0 - wLimit_);
    if (_maxW < _minW) _maxW = _minW;
  }
  if (!HasTrimSize()) SetTrimSize(ImageWidth(), ImageHeight());
  actualInversionWidth_ = ImageWidth();
  actualInversionHeight_ = ImageHeight();
  actualPixelSizeX_ = PixelSizeX();
  actualPixelSizeY_ = PixelSizeY();
  // small inversion: image at the size that samples the theoretical beam
  // by two pixels, then resample to the image size (:409-439)
  if (SmallInversion()) {
    size_t optWidth, optHeight, minWidth, minHeight;
    CalculateFFTSize(actualInversionWidth_, actualPixelSizeX_, theoreticalBeamSize_, minWidth,
                     optWidth);
    CalculateFFTSize(actualInversionHeight_, actualPixelSizeY_, theoreticalBeamSize_, minHeight,
                     optHeight);
    if (optWidth < actualInversionWidth_ || optHeight < actualInversionHeight_) {
      const size_t newWidth = std::max(std::min(optWidth, actualInversionWidth_), size_t(32));
      const size_t newHeight = std::max(std::min(optHeight, actualInversionHeight_), size_t(32));
      actualPixelSizeX_ = (double(actualInversionWidth_) * actualPixelSizeX_) / double(newWidth);
      actualPixelSizeY_ =
          (double(actualInversionHeight_) * actualPixelSizeY_) / double(newHeight);
      actualInversionWidth_ = newWidth;
      actualInversionHeight_ = newHeight;
    }
  }
  const size_t suggested = (IsFirstIteration() || wGridSize_ == 0) ? getSuggestedWGridSize() : 0;
  actualWGridSize_ = wGridSize_ != 0 ? wGridSize_ : suggested;
}

void MSGridderBase::StartMeasurementSet(const MSData& msData, bool /*isPredict*/) {
  msIndex_ = msData.msIndex;
}

template <>
void MSGridderBase::rotateVisibilities<1>(const aocommon::BandData& bandData,
                                          double shiftFactor, std::complex<float>* dataIter) {
  for (size_t ch = 0; ch != bandData.ChannelCount(); ++ch) {
    const double wShiftRad = shiftFactor / bandData.ChannelWavelength(ch);
    dataIter[ch] *= std::complex<float>(std::cos(wShiftRad), std::sin(wShiftRad));
  }
}

template <>
void MSGridderBase::readAndWeightVisibilities<1>(MSReader& msReader,
                                                 const std::vector<std::string>& /*antennaNames*/,
                                                 InversionRow& rowData,
                                                 const aocommon::BandData& curBand,
                                                 float* weightBuffer,
                                                 std::complex<float>* modelBuffer,
                                                 const bool* isSelected, GainMode /*gain_mode*/) {
  const size_t nchan = curBand.ChannelCount();
  std::complex<float>* data = rowData.data;
  const double* uvw = rowData.uvw;
  if (psfMode_ != PsfMode::kNone) {
    std::fill_n(data, nchan, std::complex<float>(1.0f, 0.0f));
    double dl = 0.0, dm = 0.0;
    if (psfMode_ == PsfMode::kSingle) {
      dl = mainImageDL_;
      dm = mainImageDM_;
    } else {
      dl = lShift_;
      dm = mShift_;
    }
    if (dl != 0.0 || dm != 0.0) {
      const double dn = std::sqrt(1.0 - dl * dl - dm * dm) - 1.0;
      const double shiftFactor = 2.0 * M_PI * (uvw[0] * dl + uvw[1] * dm + uvw[2] * dn);
      rotateVisibilities<1>(curBand, shiftFactor, data);
    }
  } else {
    msReader.ReadData(data);
  }
  if (doSubtractModel_) {
    msReader.ReadModel(modelBuffer);
    for (size_t ch = 0; ch != nchan; ++ch) data[ch] -= modelBuffer[ch];
  }
  msReader.ReadWeights(weightBuffer);
  for (size_t ch = 0; ch != nchan; ++ch)
    if (!isSelected[ch]) weightBuffer[ch] = 0.0;
  switch (visibilityWeightingMode_) {
    case VisibilityWeightingMode::Normal:
      break;
    case VisibilityWeightingMode::Squared:
      for (size_t ch = 0; ch != nchan; ++ch) weightBuffer[ch] *= weightBuffer[ch];
      break;
    case VisibilityWeightingMode::Unit:
      for (size_t ch = 0; ch != nchan; ++ch)
        if (weightBuffer[ch] != 0.0) weightBuffer[ch] = 1.0f;
      break;
  }
  scratchImageWeights_.resize(nchan);
  for (size_t ch = 0; ch != nchan; ++ch) {
    const double wl = curBand.ChannelWavelength(ch);
    scratchImageWeights_[ch] = float(GetImageWeights()->GetWeight(uvw[0] / wl, uvw[1] / wl));
  }
  for (size_t ch = 0; ch != nchan; ++ch) {
    const double cumWeight = weightBuffer[ch] * scratchImageWeights_[ch];
    if (cumWeight != 0.0) {
      _visibilityWeightSum += weightBuffer[ch];
      maxGriddedWeight_ = std::max(cumWeight, maxGriddedWeight_);
      ++griddedVisibilityCount_;
      _totalWeight += cumWeight;
    }
    weightBuffer[ch] = float(cumWeight);
    data[ch] *= float(cumWeight);
  }
}

template <>
void MSGridderBase::writeVisibilities<1>(MSProvider& msProvider,
                                         const std::vector<std::string>& /*antennaNames*/,
                                         const aocommon::BandData& /*curBand*/,
                                         std::complex<float>* buffer, GainMode /*gain_mode*/) {
  if (writerLockManager_) {
    std::unique_lock<std::mutex> guard =
        writerLockManager_->GetLock(facetGroupIndex_ * MeasurementSetCount() + msIndex_);
    msProvider.WriteModel(buffer, IsFacet());
  } else {
    msProvider.WriteModel(buffer, IsFacet());
  }
  msProvider.NextOutputRow();
}

}  // namespace wsclean_amd
