// Float image of width x height, row-major, x fastest: the part of
// aocommon::Image (external/aocommon, an un-vendored submodule of the
// reference) that the gridders use, under the same name so that a gridder
// written against the reference's MSGridderBase compiles against this mirror.
#ifndef WSCLEAN_AMD_HOST_IMAGE_H
#define WSCLEAN_AMD_HOST_IMAGE_H

#include <cstddef>
#include <vector>

namespace aocommon {

class Image {
 public:
  Image() = default;
  Image(size_t width, size_t height, float value = 0.0f)
      : width_(width), height_(height), data_(width * height, value) {}

  size_t Width() const { return width_; }
  size_t Height() const { return height_; }
  size_t Size() const { return data_.size(); }
  bool Empty() const { return data_.empty(); }
  float* Data() { return data_.data(); }
  const float* Data() const { return data_.data(); }
  float& operator[](size_t i) { return data_[i]; }
  float operator[](size_t i) const { return data_[i]; }

  // Central trimWidth x trimHeight part (aocommon::Image::Trim).
  Image Trim(size_t trimWidth, size_t trimHeight) const {
    Image out(trimWidth, trimHeight);
    const size_t sx = (width_ - trimWidth) / 2, sy = (height_ - trimHeight) / 2;
    for (size_t y = 0; y != trimHeight; ++y)
      for (size_t x = 0; x != trimWidth; ++x)
        out.data_[x + y * trimWidth] = data_[(x + sx) + (y + sy) * width_];
    return out;
  }
  // Zero-padded to width x height, keeping the centre (aocommon::Image::Untrim).
  Image Untrim(size_t width, size_t height) const {
    Image out(width, height, 0.0f);
    const size_t sx = (width - width_) / 2, sy = (height - height_) / 2;
    for (size_t y = 0; y != height_; ++y)
      for (size_t x = 0; x != width_; ++x)
        out.data_[(x + sx) + (y + sy) * width] = data_[x + y * width_];
    return out;
  }

 private:
  size_t width_ = 0, height_ = 0;
  std::vector<float> data_;
};

}  // namespace aocommon

namespace wsclean_amd {
using aocommon::Image;
}  // namespace wsclean_amd

#endif
