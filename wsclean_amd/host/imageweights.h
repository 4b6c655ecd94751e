// Gridded imaging weights (structures/imageweights.h:19-129,
// structures/imageweights.cpp:33-49, 75-144): natural, uniform and Briggs.
#ifndef WSCLEAN_AMD_HOST_IMAGEWEIGHTS_H
#define WSCLEAN_AMD_HOST_IMAGEWEIGHTS_H

#include "msprovider.h"
#include "settings.h"

#include <cmath>
#include <cstddef>
#include <vector>

namespace wsclean_amd {

class ImageWeights {
 public:
  ImageWeights(WeightMode mode, double briggsRobustness, size_t imageWidth,
               size_t imageHeight, double pixelScaleX, double pixelScaleY,
               bool weightsAsTaper = false, double superWeight = 1.0);

  double GetWeight(double u, double v) const { return sampleGridValue(u, v); }
  // Accumulates the (weight) density of every sample of a provider
  // (imageweights.cpp:75-111).
  void Grid(MSProvider& provider);
  void Grid(double u, double v, double weight) {
    int x, y;
    uvToXY(u, v, x, y);
    if (isWithinLimits(x, y)) {
      grid_[size_t(x) + size_t(y) * width_] += weight;
      totalSum_ += weight;
    }
  }
  void FinishGridding();  // imageweights.cpp:113-144
  size_t Width() const { return width_; }
  size_t Height() const { return height_; }
  const std::vector<double>& GridValues() const { return grid_; }

 private:
  void uvToXY(double u, double v, int& x, int& y) const {
    if (v < 0.0) {
      u = -u;
      v = -v;
    }
    x = int(std::floor(u * width_ * scaleX_ + width_ / 2));
    y = int(std::floor(v * height_ * scaleY_));
  }
  bool isWithinLimits(int x, int y) const {
    return x >= 0 && x < int(width_) && y < int(height_ / 2);
  }
  double sampleGridValue(double u, double v) const {
    int x, y;
    uvToXY(u, v, x, y);
    return isWithinLimits(x, y) ? grid_[size_t(x) + size_t(y) * width_] : 0.0;
  }

  WeightMode mode_;
  double robust_;
  size_t width_, height_;
  double scaleX_, scaleY_;
  std::vector<double> grid_;
  double totalSum_ = 0.0;
  bool weightsAsTaper_, finished_ = false;
};

}  // namespace wsclean_amd

#endif
