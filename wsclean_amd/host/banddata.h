// The parts of aocommon::BandData and aocommon::Polarization (external/
// aocommon, an un-vendored submodule of the reference) that the gridder
// plugin interface passes around: the channel wavelengths of one spectral
// window, a channel sub-range of it, and the correlation of a task.
#ifndef WSCLEAN_AMD_HOST_BANDDATA_H
#define WSCLEAN_AMD_HOST_BANDDATA_H

#include <algorithm>
#include <cstddef>
#include <vector>

namespace aocommon {

class BandData {
 public:
  BandData() = default;
  // Channel wavelengths in metres, in channel order.
  explicit BandData(std::vector<double> wavelengths) : wavelengths_(std::move(wavelengths)) {}
  // Channels [startChannel, endChannel) of source.
  BandData(const BandData& source, size_t startChannel, size_t endChannel)
      : wavelengths_(source.wavelengths_.begin() + startChannel,
                     source.wavelengths_.begin() + endChannel) {}

  size_t ChannelCount() const { return wavelengths_.size(); }
  double ChannelWavelength(size_t channel) const { return wavelengths_[channel]; }
  double ChannelFrequency(size_t channel) const {
    return 299792458.0 / wavelengths_[channel];
  }
  double SmallestWavelength() const {
    return wavelengths_.empty() ? 0.0
                                : *std::min_element(wavelengths_.begin(), wavelengths_.end());
  }
  double LongestWavelength() const {
    return wavelengths_.empty() ? 0.0
                                : *std::max_element(wavelengths_.begin(), wavelengths_.end());
  }
  bool operator==(const BandData& other) const { return wavelengths_ == other.wavelengths_; }

 private:
  std::vector<double> wavelengths_;
};

// The correlations a gridding task images (aocommon::Polarization values the
// w-stacking path distinguishes).
struct Polarization {
  enum PolarizationEnum { StokesI, StokesQ, StokesU, StokesV, XX, XY, YX, YY };
};
using PolarizationEnum = Polarization::PolarizationEnum;

}  // namespace aocommon

#endif
