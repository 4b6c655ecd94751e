// The gridder plugin interface of the reference, MSGridderBase
// (gridding/msgridderbase.h:44-508), with the shared host logic a gridder
// relies on: w-limits (msgridderbase.cpp:200-286), overall metadata with the
// small-inversion size (:370-446), reading and weighting of visibilities with
// the visibility counters (:515-660), the phase-centre rotation (:692-704)
// and model write-back under the writer lock (:448-493).
//
// The protected members a gridder calls have the reference's names and
// parameter lists (msgridderbase.h:232-346), so a gridder written against
// this class compiles against the reference's MSGridderBase unchanged. Facet
// beams and gain solutions are out of scope: GainMode is carried, not applied.
#ifndef WSCLEAN_AMD_HOST_MSGRIDDERBASE_H
#define WSCLEAN_AMD_HOST_MSGRIDDERBASE_H

#include "banddata.h"
#include "image.h"
#include "imageweights.h"
#include "msprovider.h"
#include "settings.h"
#include "visibilitymodifier.h"

#include <complex>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace wsclean_amd {

enum class PsfMode { kNone, kSingle, kDirectionDependent };

// scheduling/metadatacache.h:9-29
struct MetaDataCache {
  struct Entry {
    double minW, maxW, maxWWithFlags, maxBaselineUVW, maxBaselineInM, integrationTime;
  };
  std::vector<Entry> msDataVector;
};

// scheduling/writerlockmanager.h:10-34
class WriterLockManager {
 public:
  virtual ~WriterLockManager() = default;
  virtual std::unique_lock<std::mutex> GetLock(size_t writerGroupIndex) = 0;
};

class MSGridderBase {
 public:
  MSGridderBase(const Settings& settings);
  virtual ~MSGridderBase() = default;

  size_t ImageWidth() const { return imageWidth_; }
  size_t ImageHeight() const { return imageHeight_; }
  double PixelSizeX() const { return pixelSizeX_; }
  double PixelSizeY() const { return pixelSizeY_; }
  size_t ActualWGridSize() const { return actualWGridSize_; }

  void ClearMeasurementSetList() { measurementSets_.clear(); }
  void AddMeasurementSet(MSProvider* provider) { measurementSets_.push_back(provider); }
  size_t MeasurementSetCount() const { return measurementSets_.size(); }
  MSProvider& MeasurementSet(size_t i) const { return *measurementSets_[i]; }

  void SetImageWidth(size_t w) { imageWidth_ = w; }
  void SetImageHeight(size_t h) { imageHeight_ = h; }
  void SetTrimSize(size_t w, size_t h) {
    trimWidth_ = w;
    trimHeight_ = h;
  }
  size_t TrimWidth() const { return trimWidth_; }
  size_t TrimHeight() const { return trimHeight_; }
  bool HasTrimSize() const { return trimWidth_ != 0 || trimHeight_ != 0; }
  void SetIsComplex(bool c) { isComplex_ = c; }
  bool IsComplex() const { return isComplex_; }
  void SetPolarization(aocommon::PolarizationEnum p) { polarization_ = p; }
  aocommon::PolarizationEnum Polarization() const { return polarization_; }
  void SetLShift(double l) { lShift_ = l; }
  void SetMShift(double m) { mShift_ = m; }
  double LShift() const { return lShift_; }
  double MShift() const { return mShift_; }
  bool HasDenormalPhaseCentre() const { return lShift_ != 0.0 || mShift_ != 0.0; }
  void SetMainImageDL(double dl) { mainImageDL_ = dl; }
  void SetMainImageDM(double dm) { mainImageDM_ = dm; }
  GriddingKernelMode GetGridMode() const { return gridMode_; }
  void SetGridMode(GriddingKernelMode m) { gridMode_ = m; }
  void SetImageWeights(const ImageWeights* w) { imageWeights_ = w; }
  const ImageWeights* GetImageWeights() const { return imageWeights_; }
  void SetPsfMode(PsfMode m) { psfMode_ = m; }
  PsfMode GetPsfMode() const { return psfMode_; }
  void SetDoSubtractModel(bool s) { doSubtractModel_ = s; }
  void SetIsFacet(bool f) { isFacet_ = f; }
  bool IsFacet() const { return isFacet_; }
  void SetFacetGroupIndex(size_t i) { facetGroupIndex_ = i; }
  void SetWriterLockManager(WriterLockManager* m) { writerLockManager_ = m; }
  void SetIsFirstIteration(bool f) { isFirstIteration_ = f; }
  bool IsFirstIteration() const { return isFirstIteration_; }
  void SetMetaDataCache(std::unique_ptr<MetaDataCache> c) { _metaDataCache = std::move(c); }
  std::unique_ptr<MetaDataCache> AcquireMetaDataCache() { return std::move(_metaDataCache); }
  bool SmallInversion() const { return _settings.smallInversion; }

  virtual void Invert() = 0;
  virtual void Predict(std::vector<aocommon::Image>&& images) = 0;
  virtual std::vector<aocommon::Image> ResultImages() = 0;
  virtual void FreeImagingData() {}

  double StartTime() const { return startTime_; }
  double ImageWeight() const { return _totalWeight; }
  double NormalizationFactor() const { return _totalWeight; }
  double BeamSize() const { return theoreticalBeamSize_; }
  double VisibilityWeightSum() const { return _visibilityWeightSum; }
  size_t GriddedVisibilityCount() const { return griddedVisibilityCount_; }
  double MaxGriddedWeight() const { return maxGriddedWeight_; }
  double EffectiveGriddedVisibilityCount() const { return _totalWeight / maxGriddedWeight_; }

 protected:
  // gridding/msgridderbase.h:240-258
  struct MSData {
    MSProvider* msProvider = nullptr;
    size_t msIndex = 0;
    aocommon::BandData bandData;
    size_t startChannel = 0, endChannel = 0;
    size_t matchingRows = 0, totalRowsProcessed = 0;
    double minW = 0, maxW = 0, maxWWithFlags = 0, maxBaselineUVW = 0, maxBaselineInM = 0;
    size_t rowStart = 0, rowEnd = 0;
    double integrationTime = 1;
    std::vector<std::string> antennaNames;

    aocommon::BandData SelectedBand() const {
      return aocommon::BandData(bandData, startChannel, endChannel);
    }
  };

  // gridding/msgridderbase.h:260-264
  struct InversionRow {
    double uvw[3];
    size_t rowId;
    std::complex<float>* data;
  };

  size_t ActualInversionWidth() const { return actualInversionWidth_; }
  size_t ActualInversionHeight() const { return actualInversionHeight_; }
  double ActualPixelSizeX() const { return actualPixelSizeX_; }
  double ActualPixelSizeY() const { return actualPixelSizeY_; }

  // gridding/msgridderbase.cpp:164-169: the writer-lock index of the
  // measurement set written next.
  void StartMeasurementSet(const MSData& msData, bool isPredict);

  // gridding/msgridderbase.cpp:515-660: reads the row's data (or ones in PSF
  // mode) into rowData.data, subtracts the model if requested, zeroes the
  // weights of unselected samples, applies the visibility weighting mode and
  // the imaging weights and adds to the visibility counters. weightBuffer
  // returns the applied weights; modelBuffer is scratch for the model.
  template <size_t PolarizationCount>
  void readAndWeightVisibilities(MSReader& msReader, const std::vector<std::string>& antennaNames,
                                 InversionRow& rowData, const aocommon::BandData& curBand,
                                 float* weightBuffer, std::complex<float>* modelBuffer,
                                 const bool* isSelected, GainMode gain_mode);

  // gridding/msgridderbase.cpp:448-493: one row of model data written under
  // the writer lock of (facet group, measurement set).
  template <size_t PolarizationCount>
  void writeVisibilities(MSProvider& msProvider, const std::vector<std::string>& antennaNames,
                         const aocommon::BandData& curBand, std::complex<float>* buffer,
                         GainMode gain_mode);

  // gridding/msgridderbase.cpp:692-704: data[ch] *= exp(i shiftFactor / lambda_ch).
  template <size_t PolarizationCount>
  static void rotateVisibilities(const aocommon::BandData& bandData, double shiftFactor,
                                 std::complex<float>* dataIter);

  void initializeMSDataVector(std::vector<MSData>& msDataVector);

  void resetVisibilityCounters() {
    griddedVisibilityCount_ = 0;
    _totalWeight = 0.0;
    maxGriddedWeight_ = 0.0;
    _visibilityWeightSum = 0.0;
  }
  double totalWeight() const { return _totalWeight; }
  virtual size_t getSuggestedWGridSize() const = 0;

  double _maxW = 0.0, _minW = 0.0;
  const Settings& _settings;
  std::unique_ptr<MetaDataCache> _metaDataCache;
  double _visibilityWeightSum = 0.0;
  double _totalWeight = 0.0;

 private:
  void calculateWLimits(MSData& msData);
  void calculateOverallMetaData(const std::vector<MSData>& msDataVector);

  std::vector<MSProvider*> measurementSets_;
  const ImageWeights* imageWeights_ = nullptr;
  WriterLockManager* writerLockManager_ = nullptr;
  size_t imageWidth_ = 0, imageHeight_ = 0, trimWidth_ = 0, trimHeight_ = 0;
  double pixelSizeX_, pixelSizeY_;
  size_t actualInversionWidth_ = 0, actualInversionHeight_ = 0;
  double actualPixelSizeX_ = 0, actualPixelSizeY_ = 0;
  size_t wGridSize_, actualWGridSize_ = 0;
  double wLimit_;
  double lShift_ = 0.0, mShift_ = 0.0, mainImageDL_ = 0.0, mainImageDM_ = 0.0;
  bool isComplex_ = false, isFacet_ = false, doSubtractModel_ = false;
  bool isFirstIteration_ = false;
  size_t facetGroupIndex_ = 0, msIndex_ = 0;
  aocommon::PolarizationEnum polarization_ = aocommon::Polarization::StokesI;
  PsfMode psfMode_ = PsfMode::kNone;
  GriddingKernelMode gridMode_ = GriddingKernelMode::KaiserBessel;
  VisibilityWeightingMode visibilityWeightingMode_;
  double theoreticalBeamSize_ = 0.0, startTime_ = 0.0;
  size_t griddedVisibilityCount_ = 0;
  double maxGriddedWeight_ = 0.0;
  std::vector<float> scratchImageWeights_;
};

// math/calculatefftsize.h:22-32
void CalculateFFTSize(size_t size, double pixelScale, double beamSize, size_t& minimalSize,
                      size_t& optimalSize);

}  // namespace wsclean_amd

// The reference declares these in the global namespace: a gridder written
// against it uses the unqualified names.
using wsclean_amd::GriddingKernelMode;
using wsclean_amd::ImageWeights;
using wsclean_amd::MSGridderBase;
using wsclean_amd::MSProvider;
using wsclean_amd::MSReader;
using wsclean_amd::Resources;
using wsclean_amd::Settings;

#endif
