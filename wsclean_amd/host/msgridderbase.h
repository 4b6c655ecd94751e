// The gridder plugin interface of the reference, MSGridderBase
// (gridding/msgridderbase.h:44-508), with the shared host logic the
// w-stacking gridder relies on: w-limits (msgridderbase.cpp:200-286),
// overall metadata (:370-446), reading and weighting of visibilities with
// the visibility counters (:515-660) and model write-back with the writer
// lock (:448-493). Facet beams/h5parm gains and small inversion are out of
// scope for this build.
#ifndef WSCLEAN_AMD_HOST_MSGRIDDERBASE_H
#define WSCLEAN_AMD_HOST_MSGRIDDERBASE_H

#include "image.h"
#include "imageweights.h"
#include "msprovider.h"
#include "settings.h"

#include <complex>
#include <memory>
#include <mutex>
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
  void SetMetaDataCache(std::unique_ptr<MetaDataCache> c) { metaDataCache_ = std::move(c); }
  std::unique_ptr<MetaDataCache> AcquireMetaDataCache() { return std::move(metaDataCache_); }

  virtual void Invert() = 0;
  virtual void Predict(std::vector<Image>&& images) = 0;
  virtual std::vector<Image> ResultImages() = 0;
  virtual void FreeImagingData() {}

  double StartTime() const { return startTime_; }
  double ImageWeight() const { return totalWeight_; }
  double NormalizationFactor() const { return totalWeight_; }
  double BeamSize() const { return theoreticalBeamSize_; }
  double VisibilityWeightSum() const { return visibilityWeightSum_; }
  size_t GriddedVisibilityCount() const { return griddedVisibilityCount_; }
  double MaxGriddedWeight() const { return maxGriddedWeight_; }
  double EffectiveGriddedVisibilityCount() const { return totalWeight_ / maxGriddedWeight_; }

 protected:
  struct MSData {
    MSProvider* msProvider = nullptr;
    size_t msIndex = 0;
    double minW = 0, maxW = 0, maxWWithFlags = 0, maxBaselineUVW = 0, maxBaselineInM = 0;
    double integrationTime = 1;
  };

  size_t ActualInversionWidth() const { return actualInversionWidth_; }
  size_t ActualInversionHeight() const { return actualInversionHeight_; }
  double ActualPixelSizeX() const { return actualPixelSizeX_; }
  double ActualPixelSizeY() const { return actualPixelSizeY_; }

  void initializeMSDataVector(std::vector<MSData>& msDataVector);

  // readAndWeightVisibilities<1> for one row (msgridderbase.cpp:515-660):
  // reads data (or ones in PSF mode), subtracts the model if requested, zeroes
  // unselected samples, applies visibility and imaging weights, and adds to
  // the visibility counters. data/weights hold nchan values.
  void readAndWeightVisibilities(MSReader& reader, MSProvider& provider, const double* uvw,
                                 std::complex<float>* data, float* weights,
                                 const bool* isSelected);
  // writeVisibilities<1> (msgridderbase.cpp:448-493): model write under the
  // writer lock of (facet group, measurement set).
  void writeVisibilities(MSProvider& provider, size_t msIndex, const std::complex<float>* buffer);

  void resetVisibilityCounters() {
    griddedVisibilityCount_ = 0;
    totalWeight_ = 0.0;
    maxGriddedWeight_ = 0.0;
    visibilityWeightSum_ = 0.0;
  }
  double totalWeight() const { return totalWeight_; }
  virtual size_t getSuggestedWGridSize() const = 0;

  double maxW_ = 0.0, minW_ = 0.0;
  const Settings& settings_;

 private:
  void calculateWLimits(MSData& msData);
  void calculateOverallMetaData(const std::vector<MSData>& msDataVector);

  std::vector<MSProvider*> measurementSets_;
  std::unique_ptr<MetaDataCache> metaDataCache_;
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
  size_t facetGroupIndex_ = 0;
  PsfMode psfMode_ = PsfMode::kNone;
  GriddingKernelMode gridMode_ = GriddingKernelMode::KaiserBessel;
  VisibilityWeightingMode visibilityWeightingMode_;
  double theoreticalBeamSize_ = 0.0, startTime_ = 0.0;
  size_t griddedVisibilityCount_ = 0;
  double totalWeight_ = 0.0, maxGriddedWeight_ = 0.0, visibilityWeightSum_ = 0.0;
  std::vector<float> scratchImageWeights_;
};

}  // namespace wsclean_amd

#endif
