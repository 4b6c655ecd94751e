// HipWSMSGridder: the MI355X w-stacking gridder behind the MSGridderBase
// plugin interface, in the role WSMSGridder (gridding/wsmsgridder.h:20-88)
// plays for the CPU WStackingGridder<float>. It uses only the members the
// reference's MSGridderBase offers a subclass (gridding/msgridderbase.h:44-
// 508), under their names, so it compiles against that class as well as
// against this project's mirror of it. The HIP core is reached through the C
// ABI of include/hip_wstacking.h; every failure there becomes a
// std::runtime_error carrying hws_last_error().
//
// Invert reads and weights every row on the host (readAndWeightVisibilities)
// and stages the weighted rows to the device while reading; the device then
// plans the samples once and runs every w-layer pass, the small-inversion
// resampling and the trim. Predict stages the rows' uvw, untrims and
// decimates the model on the device, predicts every pass there and writes
// the rows back in order (writeVisibilities). Measurement sets whose bands
// differ go through the pass-by-pass host-pointer protocol instead.
#ifndef WSCLEAN_AMD_HOST_HIPWSMSGRIDDER_H
#define WSCLEAN_AMD_HOST_HIPWSMSGRIDDER_H

#include "msgridderbase.h"

#include "../../include/hip_wstacking.h"

#include <complex>
#include <cstddef>
#include <vector>

class HipWSMSGridder final : public MSGridderBase {
 public:
  HipWSMSGridder(const Settings& settings, const Resources& resources);
  ~HipWSMSGridder() noexcept final;

  void Invert() final;                                          // wsmsgridder.cpp:367-486
  void Predict(std::vector<aocommon::Image>&& images) final;  // wsmsgridder.cpp:488-586
  std::vector<aocommon::Image> ResultImages() final {
    if (IsComplex()) return {std::move(_realImage), std::move(_imaginaryImage)};
    return {std::move(_realImage)};
  }
  void FreeImagingData() final { releaseContext(); }

  size_t AntialiasingKernelSize() const { return _antialiasingKernelSize; }
  size_t OverSamplingFactor() const { return _overSamplingFactor; }
  size_t NPasses() const { return _nPasses; }
  // Whether the last call ran on the staged device path (all measurement
  // sets share one band) or the pass-by-pass host-pointer protocol.
  bool UsedStagedPath() const { return _stagedPath; }

 protected:
  size_t getSuggestedWGridSize() const final;  // wsmsgridder.cpp:77-128

 private:
  void makeContext();
  void releaseContext();
  void check(hws_status s) const;
  bool sameBand(const std::vector<MSData>& msDataVector) const;
  void invertStaged(std::vector<MSData>& msDataVector);
  void invertHostProtocol(std::vector<MSData>& msDataVector);
  void predictStaged(std::vector<MSData>& msDataVector, std::vector<aocommon::Image>& images);
  void predictHostProtocol(std::vector<MSData>& msDataVector,
                           std::vector<aocommon::Image>& images);

  const Resources _resources;
  const size_t _antialiasingKernelSize, _overSamplingFactor;
  const int _hipPrecision, _hipDevice, _hipTileSize;
  const double _hipMaxLayerMemory;
  hws_gridder* _context = nullptr;
  size_t _nPasses = 0;
  bool _stagedPath = false;
  aocommon::Image _realImage, _imaginaryImage;
};

#endif
