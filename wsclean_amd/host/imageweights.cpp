#include "imageweights.h"

#include <stdexcept>

namespace wsclean_amd {

ImageWeights::ImageWeights(WeightMode mode, double briggsRobustness, size_t imageWidth,
                           size_t imageHeight, double pixelScaleX, double pixelScaleY,
                           bool weightsAsTaper, double superWeight)
    : mode_(mode),
      robust_(briggsRobustness),
      width_(size_t(std::round(double(imageWidth) / superWeight))),
      height_(size_t(std::round(double(imageHeight) / superWeight))),
      scaleX_(pixelScaleX),
      scaleY_(pixelScaleY),
      weightsAsTaper_(weightsAsTaper) {
  if (width_ % 2 != 0) ++width_;
  if (height_ % 2 != 0) ++height_;
  grid_.assign(width_ * height_ / 2, 0.0);
}

void ImageWeights::Grid(MSProvider& provider) {
  if (finished_) throw std::runtime_error("Grid() called after FinishGridding()");
  const size_t nchan = provider.NChannels();
  std::vector<float> weights(nchan);
  std::unique_ptr<MSReader> reader = provider.MakeReader();
  while (reader->CurrentRowAvailable()) {
    double uInM, vInM, wInM;
    reader->ReadMeta(uInM, vInM, wInM);
    reader->ReadWeights(weights.data());
    if (weightsAsTaper_)
      for (float& w : weights)
        if (w != 0.0) w = 1.0;
    if (vInM < 0.0) {
      uInM = -uInM;
      vInM = -vInM;
    }
    for (size_t ch = 0; ch != nchan; ++ch) {
      const double wl = provider.ChannelWavelength(ch);
      Grid(uInM / wl, vInM / wl, weights[ch]);
    }
    reader->NextInputRow();
  }
}

void ImageWeights::FinishGridding() {
  if (finished_) throw std::runtime_error("FinishGridding() called twice");
  finished_ = true;
  switch (mode_) {
    case WeightMode::Briggs: {
      double avgW = 0.0;
      for (double val : grid_) avgW += val * val;
      avgW /= totalSum_;
      const double numeratorSqrt = 5.0 * std::exp((-M_LN10) * robust_);
      const double sSq = numeratorSqrt * numeratorSqrt / avgW;
      for (double& val : grid_) val = 1.0 / (1.0 + val * sSq);
    } break;
    case WeightMode::Uniform:
      for (double& val : grid_) val = (val != 0.0) ? 1.0 / val : 0.0;
      break;
    case WeightMode::Natural:
      for (double& val : grid_)
        if (val != 0.0) val = 1.0;
      break;
  }
}

}  // namespace wsclean_amd
