// The subset of the reference's Settings (main/settings.h:24-320) and
// Resources (structures/resources.h) that the w-stacking path reads.
#ifndef WSCLEAN_AMD_HOST_SETTINGS_H
#define WSCLEAN_AMD_HOST_SETTINGS_H

#include <cstddef>
#include <cstdint>

namespace wsclean_amd {

// main/settings.h:18 plus the MI355X w-stacking gridder.
enum class GridderType { WStacking, WGridder, TunedWGridder, DirectFT, IDG, HipWStacking };

// gridding/gridmode.h:6-44
enum class GriddingKernelMode {
  NearestNeighbour,
  KaiserBessel,
  KaiserBesselWithoutSinc,
  Rectangular,
  Gaussian,
  GaussianWithoutSinc,
  BlackmanNuttall,
  BlackmanNuttallWithoutSinc,
  BlackmanHarris
};

// structures/weightmode.h:12
enum class WeightMode { Natural, Uniform, Briggs };

// gridding/visibilityweightingmode.h
enum class VisibilityWeightingMode { Normal, Squared, Unit };

enum class HipPrecision { PlainFP32 = 0, Reference = 1, Mixed = 2 };

struct Settings {
  GridderType gridderType = GridderType::HipWStacking;
  GriddingKernelMode gridMode = GriddingKernelMode::KaiserBessel;
  size_t paddedImageWidth = 0, paddedImageHeight = 0;
  size_t trimmedImageWidth = 0, trimmedImageHeight = 0;
  double pixelScaleX = 0.0, pixelScaleY = 0.0;
  size_t antialiasingKernelSize = 7;
  size_t overSamplingFactor = 1023;  // main/settings.h:249
  size_t nWLayers = 0;               // 0: suggested (msgridderbase.cpp:443-445)
  double nWLayersFactor = 1.0;
  size_t widthForNWCalculation = 0, heightForNWCalculation = 0;
  double wLimit = 0.0;  // percent
  bool smallInversion = true;  // main/settings.h:267
  WeightMode weightMode = WeightMode::Uniform;
  double briggsRobustness = 0.0;
  VisibilityWeightingMode visibilityWeightingMode = VisibilityWeightingMode::Normal;
  size_t threadCount = 1;
  size_t parallelGridding = 1;
  int64_t memory = int64_t(16) << 30;  // bytes for Resources
  // MI355X gridder knobs
  HipPrecision hipPrecision = HipPrecision::Mixed;
  int hipDevice = 0;
  double hipMaxLayerMemory = 0.0;  // bytes of w-layers per pass; 0 = automatic
  size_t hipTileSize = 0;
};

class Resources {
 public:
  Resources() = default;
  Resources(size_t nCpus, int64_t memory) : nCpus_(nCpus), memory_(memory) {}
  size_t NCpus() const { return nCpus_; }
  int64_t Memory() const { return memory_; }
  // structures/resources.cpp:16-22
  Resources GetPart(size_t partSize) const {
    const size_t n = (nCpus_ + partSize - 1) / partSize;
    return Resources(n == 0 ? 1 : n, memory_ / int64_t(partSize));
  }

 private:
  size_t nCpus_ = 1;
  int64_t memory_ = 0;
};

}  // namespace wsclean_amd

#endif
