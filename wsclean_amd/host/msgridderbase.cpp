#include "msgridderbase.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace wsclean_amd {

MSGridderBase::MSGridderBase(const Settings& settings)
    : settings_(settings),
      pixelSizeX_(settings.pixelScaleX),
      pixelSizeY_(settings.pixelScaleY),
      wGridSize_(settings.nWLayers),
      wLimit_(settings.wLimit / 100.0),
      visibilityWeightingMode_(settings.visibilityWeightingMode) {}

// gridding/msgridderbase.cpp:200-286 for a single-correlation provider.
void MSGridderBase::calculateWLimits(MSData& msData) {
  MSProvider& provider = *msData.msProvider;
  const size_t nchan = provider.NChannels();
  msData.maxW = 0.0;
  msData.maxWWithFlags = 0.0;
  msData.minW = 1e100;
  msData.maxBaselineUVW = 0.0;
  msData.maxBaselineInM = 0.0;
  std::vector<float> weights(nchan);
  double smallestWl = std::numeric_limits<double>::max(), longestWl = 0.0;
  for (size_t ch = 0; ch != nchan; ++ch) {
    smallestWl = std::min(smallestWl, provider.ChannelWavelength(ch));
    longestWl = std::max(longestWl, provider.ChannelWavelength(ch));
  }
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
        const double wl = provider.ChannelWavelength(ch);
        const double wInL = meta.wInM / wl;
        msData.maxWWithFlags = std::max(msData.maxWWithFlags, std::fabs(wInL));
        if (weights[ch] != 0.0) {
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
  if (!metaDataCache_) metaDataCache_.reset(new MetaDataCache());
  const bool hasCache = !metaDataCache_->msDataVector.empty();
  if (!hasCache) metaDataCache_->msDataVector.resize(MeasurementSetCount());
  startTime_ = 0.0;
  for (size_t i = 0; i != MeasurementSetCount(); ++i) {
    MSData& d = msDataVector[i];
    d.msIndex = i;
    d.msProvider = measurementSets_[i];
    startTime_ = (i == 0) ? d.msProvider->StartTime() : std::min(startTime_, d.msProvider->StartTime());
    MetaDataCache::Entry& e = metaDataCache_->msDataVector[i];
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

// gridding/msgridderbase.cpp:370-446 (small inversion not available)
void MSGridderBase::calculateOverallMetaData(const std::vector<MSData>& msDataVector) {
  maxW_ = 0.0;
  minW_ = std::numeric_limits<double>::max();
  double maxBaseline = 0.0;
  for (const MSData& d : msDataVector) {
    maxBaseline = std::max(maxBaseline, d.maxBaselineUVW);
    maxW_ = std::max(maxW_, d.maxW);
    minW_ = std::min(minW_, d.minW);
  }
  if (minW_ > maxW_) minW_ = maxW_;
  theoreticalBeamSize_ = 1.0 / maxBaseline;
  if (wLimit_ != 0.0) {
    maxW_ *= (1.0 - wLimit_);
    if (maxW_ < minW_) maxW_ = minW_;
  }
  if (!HasTrimSize()) SetTrimSize(ImageWidth(), ImageHeight());
  actualInversionWidth_ = ImageWidth();
  actualInversionHeight_ = ImageHeight();
  actualPixelSizeX_ = PixelSizeX();
  actualPixelSizeY_ = PixelSizeY();
  const size_t suggested = (IsFirstIteration() || wGridSize_ == 0) ? getSuggestedWGridSize() : 0;
  actualWGridSize_ = wGridSize_ != 0 ? wGridSize_ : suggested;
}

void MSGridderBase::readAndWeightVisibilities(MSReader& reader, MSProvider& provider,
                                              const double* uvw, std::complex<float>* data,
                                              float* weights, const bool* isSelected) {
  const size_t nchan = provider.NChannels();
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
      for (size_t ch = 0; ch != nchan; ++ch) {
        const double rad = shiftFactor / provider.ChannelWavelength(ch);
        data[ch] *= std::complex<float>(std::cos(rad), std::sin(rad));
      }
    }
  } else {
    reader.ReadData(data);
  }
  if (doSubtractModel_) {
    std::vector<std::complex<float>> model(nchan);
    reader.ReadModel(model.data());
    for (size_t ch = 0; ch != nchan; ++ch) data[ch] -= model[ch];
  }
  reader.ReadWeights(weights);
  for (size_t ch = 0; ch != nchan; ++ch)
    if (!isSelected[ch]) weights[ch] = 0.0;
  switch (visibilityWeightingMode_) {
    case VisibilityWeightingMode::Normal:
      break;
    case VisibilityWeightingMode::Squared:
      for (size_t ch = 0; ch != nchan; ++ch) weights[ch] *= weights[ch];
      break;
    case VisibilityWeightingMode::Unit:
      for (size_t ch = 0; ch != nchan; ++ch)
        if (weights[ch] != 0.0) weights[ch] = 1.0f;
      break;
  }
  scratchImageWeights_.resize(nchan);
  for (size_t ch = 0; ch != nchan; ++ch) {
    const double wl = provider.ChannelWavelength(ch);
    scratchImageWeights_[ch] = float(GetImageWeights()->GetWeight(uvw[0] / wl, uvw[1] / wl));
  }
  for (size_t ch = 0; ch != nchan; ++ch) {
    const double cumWeight = weights[ch] * scratchImageWeights_[ch];
    if (cumWeight != 0.0) {
      visibilityWeightSum_ += weights[ch];
      maxGriddedWeight_ = std::max(cumWeight, maxGriddedWeight_);
      ++griddedVisibilityCount_;
      totalWeight_ += cumWeight;
    }
    weights[ch] = float(cumWeight);
    data[ch] *= float(cumWeight);
  }
}

void MSGridderBase::writeVisibilities(MSProvider& provider, size_t msIndex,
                                      const std::complex<float>* buffer) {
  if (writerLockManager_) {
    std::unique_lock<std::mutex> guard =
        writerLockManager_->GetLock(facetGroupIndex_ * MeasurementSetCount() + msIndex);
    provider.WriteModel(buffer, IsFacet());
  } else {
    provider.WriteModel(buffer, IsFacet());
  }
  provider.NextOutputRow();
}

}  // namespace wsclean_amd
